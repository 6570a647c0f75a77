"""NumPy restatement of diffmpm's explicit MPM step (oracle, test infrastructure).

Every function cites the reference file:line whose arithmetic it restates.
Shapes follow the reference but drop its singleton axes: particle vectors are
``(N, ndim)``, tensors ``(N, 6)`` in Voigt order [xx, yy, zz, xy, yz, xz],
scalars ``(N,)``; nodal vectors ``(M, ndim)``, nodal mass ``(M,)``.

Index semantics follow JAX, which the reference relies on: gathers clamp
out-of-range indices, scatters drop them, and negative indices are first
normalised by adding the axis length.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import List, Optional, Sequence, Tuple

import numpy as np

SNAP = 1e-12  # element.py:186,199,321,326; particle.py:311


# ----------------------------------------------------------------------------
# JAX indexing semantics
# ----------------------------------------------------------------------------
def _normalise(idx, size):
    idx = np.asarray(idx, dtype=np.int64)
    return np.where(idx < 0, idx + size, idx)


def gather(arr, idx):
    """``arr[idx]`` with JAX's clamping of out-of-range indices."""
    i = np.clip(_normalise(idx, arr.shape[0]), 0, arr.shape[0] - 1)
    return arr[i]


def scatter_add(arr, idx, vals):
    """``arr.at[idx].add(vals)`` with JAX's drop of out-of-range indices.

    ``np.add.at`` applies the updates in order, like the reference's serial
    ``lax.fori_loop`` over particles (element.py:138-152).
    """
    i = _normalise(idx, arr.shape[0]).reshape(-1)
    v = np.asarray(vals, dtype=arr.dtype).reshape((i.shape[0],) + arr.shape[1:])
    ok = (i >= 0) & (i < arr.shape[0])
    np.add.at(arr, i[ok], v[ok])


def lu_inverse(J):
    """Batched inverse of 1x1 / 2x2 matrices by LU with partial pivoting (as
    ``jnp.linalg.inv`` computes it); singular matrices give inf/NaN entries
    instead of raising, like JAX."""
    J = np.asarray(J)
    out = np.empty_like(J)
    with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
        if J.shape[-1] == 1:
            out[..., 0, 0] = 1 / J[..., 0, 0]
            return out
        a, b, c, d = J[..., 0, 0], J[..., 0, 1], J[..., 1, 0], J[..., 1, 1]
        swap = np.abs(c) > np.abs(a)
        p, q = np.where(swap, c, a), np.where(swap, d, b)   # pivot row
        r, s = np.where(swap, a, c), np.where(swap, b, d)   # other row
        lf = r / p
        u22 = s - lf * q
        # columns of the inverse: solve for e0 and e1 (rows permuted if swap)
        for col in range(2):
            e_p = np.where(swap, 1.0 if col == 1 else 0.0, 1.0 if col == 0 else 0.0)
            e_r = np.where(swap, 1.0 if col == 0 else 0.0, 1.0 if col == 1 else 0.0)
            y1 = e_r - lf * e_p
            x1 = y1 / u22
            x0 = (e_p - q * x1) / p
            out[..., 0, col] = x0
            out[..., 1, col] = x1
    return out


def snap(x):
    """``where(|x| < 1e-12, 0, x)`` (element.py:185-189)."""
    return np.where(np.abs(x) < SNAP, np.zeros_like(x), x)


def nan_to_num(x):
    """``jnp.nan_to_num`` (element.py:311): NaN -> 0, +-inf -> +-max finite."""
    return np.nan_to_num(x, nan=0.0, posinf=np.finfo(x.dtype).max,
                         neginf=-np.finfo(x.dtype).max)


# ----------------------------------------------------------------------------
# Time functions (functions.py)
# ----------------------------------------------------------------------------
class Unit:
    """functions.py:18-22: value(t) = 1."""

    def __init__(self, id=-1):
        self.id = id

    def value(self, t, dtype=np.float32):
        return 1.0


class Linear:
    """functions.py:34-41: ``jnp.interp(t, xvalues, fxvalues)``.

    Outside [x0, xn] the end values are held (np.interp semantics).
    """

    def __init__(self, id, xvalues, fxvalues):
        self.id = id
        self.xvalues = np.asarray(xvalues, dtype=np.float64)
        self.fxvalues = np.asarray(fxvalues, dtype=np.float64)

    def value(self, t, dtype=np.float32):
        xp = self.xvalues.astype(dtype)
        fp = self.fxvalues.astype(dtype)
        t = dtype(t)
        n = xp.shape[0]
        i = int(np.clip(np.searchsorted(xp, t, side="right"), 1, n - 1))
        dx = xp[i] - xp[i - 1]
        if dx == 0:
            f = fp[i - 1]
        else:
            f = fp[i - 1] + ((t - xp[i - 1]) / dx) * (fp[i] - fp[i - 1])
        if t < xp[0]:
            f = fp[0]
        if t > xp[-1]:
            f = fp[-1]
        return dtype(f)


# ----------------------------------------------------------------------------
# Materials (material.py)
# ----------------------------------------------------------------------------
class LinearElastic:
    """material.py:54-116.  Props are combined in Python double precision and
    the 6x6 De is rounded to the working dtype (``jnp.array`` of floats)."""

    _props = ("density", "youngs_modulus", "poisson_ratio")

    def __init__(self, props):
        for k in self._props:  # material.py:44-51
            if k not in props:
                raise KeyError(k)
        self.properties = dict(props)
        E = props["youngs_modulus"]
        nu = props["poisson_ratio"]
        K = E / (3 * (1 - 2 * nu))                     # material.py:72
        G = E / (2 * (1 + nu))                         # material.py:95-97
        a1 = K + (4 * G / 3)                           # material.py:99
        a2 = K - (2 * G / 3)                           # material.py:100
        self.de64 = np.array([[a1, a2, a2, 0, 0, 0],
                              [a2, a1, a2, 0, 0, 0],
                              [a2, a2, a1, 0, 0, 0],
                              [0, 0, 0, G, 0, 0],
                              [0, 0, 0, 0, G, 0],
                              [0, 0, 0, 0, 0, G]], dtype=np.float64)

    def compute_stress(self, dstrain):
        """material.py:113-116: dsigma = De @ dstrain."""
        de = self.de64.astype(dstrain.dtype)
        return np.einsum("kl,nl->nk", de, dstrain)


class SimpleMaterial:
    """material.py:119-131: dsigma = E * dstrain."""

    _props = ("E", "density")

    def __init__(self, props):
        for k in self._props:
            if k not in props:
                raise KeyError(k)
        self.properties = dict(props)

    def compute_stress(self, dstrain):
        return dstrain * dstrain.dtype.type(self.properties["E"])


# ----------------------------------------------------------------------------
# Elements + nodes (element.py, node.py, constraint.py)
# ----------------------------------------------------------------------------
@dataclass
class Constraint:
    """constraint.py:4-44."""
    dir: int
    velocity: float


@dataclass
class NodalForce:
    """forces.py:9-18."""
    node_ids: np.ndarray
    function: object
    dir: int
    force: float


@dataclass
class ParticleTraction:
    """forces.py:26-41."""
    pset: Sequence[int]
    pids: np.ndarray
    function: object
    dir: int
    traction: float


class Grid:
    """node.py:34-82: nodal state.  ``mass`` starts at 1 (node.py:67)."""

    def __init__(self, loc):
        self.loc = loc
        dt = loc.dtype
        self.velocity = np.zeros_like(loc)
        self.acceleration = np.zeros_like(loc)
        self.mass = np.ones(loc.shape[0], dtype=dt)
        self.momentum = np.zeros_like(loc)
        self.f_int = np.zeros_like(loc)
        self.f_ext = np.zeros_like(loc)
        self.f_damp = np.zeros_like(loc)

    def total_force(self):
        """node.py:123-125."""
        return self.f_int + self.f_ext + self.f_damp


class Elements:
    """Linear1D (element.py:343-600) and Quadrilateral4Node (element.py:603-909)."""

    def __init__(self, kind, nelements, el_len, constraints=(),
                 nodal_forces=(), dtype=np.float32):
        assert kind in ("linear1d", "quad4")
        self.kind = kind
        self.dtype = dtype
        self.ndim = 1 if kind == "linear1d" else 2
        self.constraints: List[Tuple[np.ndarray, Constraint]] = list(constraints)
        self.nodal_forces: List[NodalForce] = list(nodal_forces)
        if kind == "linear1d":
            self.nx = int(nelements)
            self.ny = 1
            self.hx = dtype(el_len)
            self.total_elements = self.nx
            # element.py:397-401: arange(n+1) * el_len, FP32 product
            loc = (np.arange(self.nx + 1).astype(dtype) * self.hx).reshape(-1, 1)
        else:
            self.nx, self.ny = (int(v) for v in nelements)
            h = np.broadcast_to(np.asarray(el_len, dtype=np.float64), (2,))
            self.hx, self.hy = dtype(h[0]), dtype(h[1])
            self.total_elements = self.nx * self.ny
            # element.py:667-680: node n = j*(nx+1) + i at (i*hx, j*hy)
            ii, jj = np.meshgrid(np.arange(self.nx + 1), np.arange(self.ny + 1))
            loc = np.stack([ii.reshape(-1).astype(dtype) * self.hx,
                            jj.reshape(-1).astype(dtype) * self.hy], axis=1)
        self.nodes = Grid(loc.astype(dtype))
        self.volume = np.ones(self.total_elements, dtype=dtype)

    # -- connectivity -------------------------------------------------------
    def id_to_node_ids(self, eid):
        """element.py:414-428 (1D) and 692-723 (Q4); eid may be an array."""
        eid = np.asarray(eid, dtype=np.int64)
        if self.kind == "linear1d":
            return np.stack([eid, eid + 1], axis=-1)
        nx = self.nx
        ll = (eid // nx) * (nx + 1) + eid % nx
        return np.stack([ll, ll + 1, ll + nx + 2, ll + nx + 1], axis=-1)

    # -- shape functions ----------------------------------------------------
    def shapefn(self, xi):
        """element.py:430-454 (1D), 725-757 (Q4).  xi (N, ndim) -> (N, nn)."""
        xi = np.asarray(xi)
        if self.kind == "linear1d":
            x = xi[:, 0]
            return np.stack([0.5 * (1 - x), 0.5 * (1 + x)], axis=1)
        x, y = xi[:, 0], xi[:, 1]
        return np.stack([0.25 * (1 - x) * (1 - y), 0.25 * (1 + x) * (1 - y),
                         0.25 * (1 + x) * (1 + y), 0.25 * (1 - x) * (1 + y)],
                        axis=1)

    def shapefn_natural_grad(self, xi):
        """element.py:456-486 (1D: jacobian of shapefn), 759-789 (Q4).

        xi (N, ndim) -> (N, nn, ndim)."""
        xi = np.asarray(xi)
        n = xi.shape[0]
        if self.kind == "linear1d":
            g = np.empty((n, 2, 1), dtype=xi.dtype)
            g[:, 0, 0] = -0.5
            g[:, 1, 0] = 0.5
            return g
        x, y = xi[:, 0], xi[:, 1]
        g = np.empty((n, 4, 2), dtype=xi.dtype)
        g[:, 0, 0] = -0.25 * (1 - y); g[:, 0, 1] = -0.25 * (1 - x)
        g[:, 1, 0] = 0.25 * (1 - y);  g[:, 1, 1] = -0.25 * (1 + x)
        g[:, 2, 0] = 0.25 * (1 + y);  g[:, 2, 1] = 0.25 * (1 + x)
        g[:, 3, 0] = -0.25 * (1 + y); g[:, 3, 1] = 0.25 * (1 - x)
        return g

    def shapefn_grad(self, xi, coords):
        """element.py:488-515 (1D), 791-818 (Q4):
        J = dN/dxi^T @ X_e;  dN/dx = dN/dxi @ inv(J)^T (LU inverse)."""
        g = self.shapefn_natural_grad(xi)
        J = np.einsum("nka,nkb->nab", g, coords)
        Jinv = lu_inverse(J)
        return np.einsum("nka,nca->nkc", g, Jinv)

    # -- element search (A1) -------------------------------------------------
    def element_ids(self, loc):
        """Closed form of the reference's element search.

        Q4 (element.py:820-840): lower_left = max{n : X_n <= x and Y_n <= y}
        (or -1), eid = lower_left - lower_left // (nx+1).  With I = max{i :
        X_i <= x} and J = max{j : Y_j <= y} this is J*nx + I (so x >= X_nx
        wraps into the next cell row, y >= Y_ny gives eid >= nx*ny), and 0 when
        either set is empty (points left of / below the mesh, NaN).
        1D (element.py:517-541): eid = I when X_I <= x < X_{I+1}, else -1.
        """
        loc = np.asarray(loc)

        def last_le(coords, v):
            # max{i : coords_i <= v} for nondecreasing coords, -1 if none or NaN
            i = np.searchsorted(coords, v, side="right") - 1
            return np.where(np.isnan(v), -1, i)

        I = last_le(self.nodes.loc[: self.nx + 1, 0], loc[:, 0])
        if self.kind == "linear1d":
            # idl: largest n with X_n <= x, or M if none; idg: first n with
            # X_n > x minus 1, or -2 if none; eid = idl if idl == idg else -1,
            # which is I for 0 <= I < nx
            return np.where((I >= 0) & (I < self.nx), I, -1).astype(np.int32)
        J = last_le(self.nodes.loc[:: self.nx + 1, 1], loc[:, 1])
        eid = np.where((I >= 0) & (J >= 0), J * self.nx + I, 0)
        return eid.astype(np.int32)

    def element_ids_bruteforce(self, loc):
        """Literal restatement of element.py:827-837 for tiny meshes: the
        nonzero/isin set construction, O(N*M^2)."""
        out = []
        M = self.nodes.loc.shape[0]
        for p in range(loc.shape[0]):
            if self.kind == "linear1d":
                x = loc[p, 0]
                le = [n for n in range(M) if self.nodes.loc[n, 0] <= x]
                idl = le[-1] if le else M
                gt = [n for n in range(M) if self.nodes.loc[n, 0] > x]
                idg = (gt[0] if gt else -1) - 1
                out.append(idl if idl == idg else -1)
                continue
            xidl = [n for n in range(M) if self.nodes.loc[n, 0] <= loc[p, 0]]
            yidl = [n for n in range(M) if self.nodes.loc[n, 1] <= loc[p, 1]]
            xidl = xidl + [-1] * (M - len(xidl))
            yidl = yidl + [-1] * (M - len(yidl))
            ys = set(yidl)
            ll = max(v if v in ys else -1 for v in xidl)
            out.append(ll - ll // (self.nx + 1))
        return np.asarray(out, dtype=np.int32)

    # -- element volume -------------------------------------------------------
    def compute_volume(self):
        """Q4: element.py:903-909 (Bretschneider's formula in the working
        dtype, = hx*hy).  1D: the intended X_{e+1} - X_e; the reference's
        element.py:545-546 broadcasts to (E,1,E) and crashes for E > 1."""
        dt = self.dtype
        if self.kind == "linear1d":
            X = self.nodes.loc[:, 0]
            self.volume = (X[1:] - X[:-1]).astype(dt)
            return
        a = c = self.hy
        b = d = self.hx
        p = q = np.sqrt(a ** 2 + b ** 2)
        vol = dt(0.25) * np.sqrt(dt(4) * p * p * q * q
                                 - (a * a + c * c - b * b - d * d) ** 2)
        self.volume = np.full(self.total_elements, vol, dtype=dt)

    # -- P2G (A5, A6, A11, A12, A13, A15) ------------------------------------
    def _mapped(self, parts):
        return (self.shapefn(parts.reference_loc),
                self.id_to_node_ids(parts.element_ids))

    def compute_nodal_mass(self, parts):
        """element.py:122-152."""
        self.nodes.mass[:] = 0
        N, ids = self._mapped(parts)
        scatter_add(self.nodes.mass, ids, parts.mass[:, None] * N)

    def compute_nodal_momentum(self, parts):
        """element.py:154-189 (snap after the scatter)."""
        self.nodes.momentum[:] = 0
        N, ids = self._mapped(parts)
        pmom = parts.mass[:, None] * parts.velocity
        scatter_add(self.nodes.momentum, ids, N[:, :, None] * pmom[:, None, :])
        self.nodes.momentum = snap(self.nodes.momentum)

    def compute_velocity(self, parts):
        """element.py:191-202: keep the old velocity where m == 0."""
        m = self.nodes.mass[:, None]
        with np.errstate(divide="ignore", invalid="ignore"):
            v = np.where(m == 0, self.nodes.velocity, self.nodes.momentum / m)
        self.nodes.velocity = snap(v)

    def compute_external_force(self, parts):
        """element.py:204-234 (zero, then sum N * f_ext_p)."""
        self.nodes.f_ext[:] = 0
        N, ids = self._mapped(parts)
        scatter_add(self.nodes.f_ext, ids, N[:, :, None] * parts.f_ext[:, None, :])

    def compute_body_force(self, parts, gravity):
        """element.py:236-268 (accumulates sum N * (m_p g))."""
        N, ids = self._mapped(parts)
        g = np.broadcast_to(np.asarray(gravity), (self.ndim,))
        mg = parts.mass[:, None] * g.astype(self.dtype)[None, :]
        scatter_add(self.nodes.f_ext, ids, N[:, :, None] * mg[:, None, :])

    def apply_concentrated_nodal_forces(self, parts, t):
        """element.py:270-284: f_ext[ids, dir] += f(t) * F."""
        for cnf in self.nodal_forces:
            factor = cnf.function.value(t, self.dtype)
            ids = _normalise(np.asarray(cnf.node_ids), self.nodes.f_ext.shape[0])
            col = self.nodes.f_ext[:, cnf.dir].copy()
            np.add.at(col, ids, self.dtype(factor * cnf.force))
            self.nodes.f_ext[:, cnf.dir] = col

    def apply_particle_traction_forces(self, parts):
        """element.py:286-303 (accumulates sum N * traction_p)."""
        N, ids = self._mapped(parts)
        scatter_add(self.nodes.f_ext, ids, N[:, :, None] * parts.traction[:, None, :])

    def compute_internal_force(self, parts):
        """1D element.py:548-600: f_int = -sum V sigma_xx dN/dx.
        Q4 element.py:842-901: f_int = -sum V [dNx sxx + dNy sxy,
                                               dNy syy + dNx sxy]."""
        self.nodes.f_int[:] = 0
        ids = self.id_to_node_ids(parts.element_ids)
        coords = gather(self.nodes.loc, ids)
        dN = self.shapefn_grad(parts.reference_loc, coords)
        nV = -parts.volume[:, None]
        s = parts.stress
        if self.kind == "linear1d":
            upd = (nV * s[:, 0:1] * dN[:, :, 0])[:, :, None]
        else:
            fx = dN[:, :, 0] * s[:, 0:1] + dN[:, :, 1] * s[:, 3:4]
            fy = dN[:, :, 1] * s[:, 1:2] + dN[:, :, 0] * s[:, 3:4]
            upd = np.stack([nV * fx, nV * fy], axis=2)
        scatter_add(self.nodes.f_int, ids, upd)

    # -- grid update (A7, A16) -------------------------------------------------
    def apply_boundary_constraints(self, *args):
        """element.py:331-334 -> constraint.py:29-44, in list order."""
        nodes = self.nodes
        for ids, c in self.constraints:
            ids = np.asarray(ids)
            nodes.velocity[ids, c.dir] = c.velocity
            nodes.momentum[ids, c.dir] = nodes.mass[ids] * self.dtype(c.velocity)
            nodes.acceleration[ids, c.dir] = 0

    def update_nodal_acceleration_velocity(self, parts, dt):
        """element.py:305-329."""
        nodes = self.nodes
        with np.errstate(divide="ignore", invalid="ignore"):
            a = nan_to_num(nodes.total_force() / nodes.mass[:, None])
        nodes.acceleration = a
        nodes.velocity = nodes.velocity + a * self.dtype(dt)
        self.apply_boundary_constraints()
        nodes.momentum = nodes.mass[:, None] * nodes.velocity
        nodes.velocity = snap(nodes.velocity)
        nodes.acceleration = snap(nodes.acceleration)


# ----------------------------------------------------------------------------
# Particles (particle.py)
# ----------------------------------------------------------------------------
class Particles:
    """particle.py:12-91 (fresh state, ``initialized is None`` branch)."""

    def __init__(self, loc, material, element_ids=None, dtype=np.float32):
        loc = np.asarray(loc, dtype=dtype)
        if loc.ndim == 3:
            loc = loc[:, 0, :]
        n, nd = loc.shape
        self.dtype = dtype
        self.material = material
        self.loc = loc.copy()
        self.element_ids = (np.zeros(n, np.int32) if element_ids is None
                            else np.asarray(element_ids, np.int32))
        self.mass = np.ones(n, dtype)
        self.density = np.full(n, material.properties["density"], dtype)
        self.volume = np.ones(n, dtype)
        self.size = np.zeros((n, nd), dtype)
        self.velocity = np.zeros((n, nd), dtype)
        self.acceleration = np.zeros((n, nd), dtype)
        self.momentum = np.zeros((n, nd), dtype)
        self.strain = np.zeros((n, 6), dtype)
        self.stress = np.zeros((n, 6), dtype)
        self.strain_rate = np.zeros((n, 6), dtype)
        self.dstrain = np.zeros((n, 6), dtype)
        self.f_ext = np.zeros((n, nd), dtype)
        self.traction = np.zeros((n, nd), dtype)
        self.reference_loc = np.zeros((n, nd), dtype)
        self.dvolumetric_strain = np.zeros(n, dtype)
        self.volumetric_strain_centroid = np.zeros(n, dtype)
        # Optional per-particle Young's modulus (an extension; the reference
        # has scalar material properties only).
        self.youngs_modulus = None

    def __len__(self):
        return self.loc.shape[0]

    def set_particle_element_ids(self, el: Elements):
        """element.py:517-541 / 820-840."""
        self.element_ids = el.element_ids(self.loc)

    def compute_volume(self, el: Elements, total_elements=None):
        """particle.py:158-178: V_p = V_e / count(e), size = V^(1/ndim),
        m = V * rho."""
        eid = self.element_ids
        counts = np.bincount(np.clip(eid, 0, None).astype(np.int64),
                             minlength=el.total_elements)[: el.total_elements]
        counts = counts.astype(self.dtype)
        vol = gather(el.volume, eid) / gather(counts, eid)
        self.volume = vol.astype(self.dtype)
        nd = self.loc.shape[1]
        sz = np.power(self.volume, self.dtype(1.0 / nd))
        self.size = np.repeat(sz[:, None], nd, axis=1).astype(self.dtype)
        self.mass = (vol * self.density).astype(self.dtype)

    def update_natural_coords(self, el: Elements):
        """particle.py:180-206: xi = (x - (X0 + X2)/2) * (2/(X2 - X0)); for
        the 2-node element index 2 clamps to node 1."""
        t = gather(el.nodes.loc, el.id_to_node_ids(self.element_ids))
        t0 = t[:, 0, :]
        t2 = t[:, min(2, t.shape[1] - 1), :]
        with np.errstate(divide="ignore", invalid="ignore"):
            self.reference_loc = (self.loc - (t0 + t2) / 2) * (2 / (t2 - t0))

    def update_position_velocity(self, el: Elements, dt, velocity_update):
        """particle.py:208-247 (G2P)."""
        N = el.shapefn(self.reference_loc)
        ids = el.id_to_node_ids(self.element_ids)
        nv = np.sum(N[:, :, None] * gather(el.nodes.velocity, ids), axis=1)
        na = np.sum(N[:, :, None] * gather(el.nodes.acceleration, ids), axis=1)
        dt = self.dtype(dt)
        if velocity_update:
            self.velocity = nv
        else:
            self.velocity = self.velocity + na * dt
        self.loc = self.loc + nv * dt
        self.momentum = self.mass[:, None] * self.velocity

    def _strain_rate(self, dN, el: Elements):
        """particle.py:281-313.  L = dN^T @ v_e; rate = [L00, L11, 0,
        L01 + L10, 0, 0], snapped.  For the 1D element L is 1x1 and the
        out-of-range indices clamp, giving [L, L, 0, 2L, 0, 0]."""
        v = gather(el.nodes.velocity, el.id_to_node_ids(self.element_ids))
        L = np.einsum("nka,nkb->nab", dN, v)
        n = dN.shape[0]
        r = np.zeros((n, 6), self.dtype)
        if L.shape[1] == 1:
            r[:, 0] = L[:, 0, 0]
            r[:, 1] = L[:, 0, 0]
            r[:, 3] = L[:, 0, 0] + L[:, 0, 0]
        else:
            r[:, 0] = L[:, 0, 0]
            r[:, 1] = L[:, 1, 1]
            r[:, 3] = L[:, 0, 1] + L[:, 1, 0]
        return snap(r)

    def compute_strain(self, el: Elements, dt):
        """particle.py:249-279 (rate at xi, then centroid rate at xi = 0)."""
        dt = self.dtype(dt)
        coords = gather(el.nodes.loc, el.id_to_node_ids(self.element_ids))
        dN = el.shapefn_grad(self.reference_loc, coords)
        self.strain_rate = self._strain_rate(dN, el)
        self.dstrain = self.strain_rate * dt
        self.strain = self.strain + self.dstrain
        dNc = el.shapefn_grad(np.zeros_like(self.loc), coords)
        rc = self._strain_rate(dNc, el)
        nd = self.loc.shape[1]
        self.dvolumetric_strain = dt * rc[:, :nd].sum(axis=1)
        self.volumetric_strain_centroid = (self.volumetric_strain_centroid
                                           + self.dvolumetric_strain)

    def compute_stress(self, *args):
        """particle.py:315-322 (+ per-particle E extension)."""
        if self.youngs_modulus is not None and isinstance(self.material, SimpleMaterial):
            # per-particle E in place of the material's E
            self.stress = self.stress + self.dstrain * self.youngs_modulus[:, None]
            return
        if self.youngs_modulus is not None:
            E = self.youngs_modulus
            nu = self.material.properties["poisson_ratio"]
            c = LinearElastic({"density": 1.0, "youngs_modulus": 1.0,
                               "poisson_ratio": nu}).de64.astype(self.dtype)
            # coefficients scaled per particle, then De_p @ dstrain
            dsig = np.einsum("nkl,nl->nk", c[None, :, :] * E[:, None, None], self.dstrain)
            self.stress = self.stress + dsig
            return
        self.stress = self.stress + self.material.compute_stress(self.dstrain)

    def update_volume(self, *args):
        """particle.py:324-327."""
        f = 1 + self.dvolumetric_strain
        self.volume = self.volume * f
        self.density = self.density / f

    def assign_traction(self, pids, dir, traction_):
        """particle.py:329-343: traction[pids, dir] += t * V / size[dir]."""
        pids = np.asarray(pids)
        col = self.traction[:, dir].copy()
        val = (self.dtype(traction_) * self.volume[pids]) / self.size[pids, dir]
        np.add.at(col, pids, val)
        self.traction[:, dir] = col

    def zero_traction(self, *args):
        """particle.py:345-347."""
        self.traction[:] = 0


# ----------------------------------------------------------------------------
# Mesh dispatch, schemes and solver (mesh.py, scheme.py, solver.py)
# ----------------------------------------------------------------------------
class Mesh:
    """mesh.py:13-126: method-major fan-out over particle sets."""

    def __init__(self, elements: Elements, particles: Sequence[Particles],
                 particle_surface_traction: Sequence[ParticleTraction] = ()):
        self.elements = elements
        self.particles = list(particles)
        self.particle_tractions = list(particle_surface_traction)

    def apply_on_elements(self, name, args=()):
        """mesh.py:31-44."""
        f = getattr(self.elements, name)
        for ps in self.particles:
            f(ps, *args)

    def apply_on_particles(self, name, args=()):
        """mesh.py:46-58."""
        for ps in self.particles:
            getattr(ps, name)(self.elements, *args)

    def apply_traction_on_particles(self, t):
        """mesh.py:60-77."""
        self.apply_on_particles("zero_traction")
        for pt in self.particle_tractions:
            val = self.elements.dtype(pt.function.value(t, self.elements.dtype)
                                      * pt.traction)
            for pset_id in pt.pset:
                self.particles[pset_id].assign_traction(pt.pids, pt.dir, val)
        self.apply_on_elements("apply_particle_traction_forces")


class MPMExplicit:
    """solver.py:18-85 (init) and 159-221 (``solve_jit``); scheme.py phases."""

    def __init__(self, mesh: Mesh, dt, scheme="usf", velocity_update=False,
                 sim_steps=1):
        if scheme not in ("usf", "usl"):
            raise ValueError(f"Please select scheme from ('usf', 'usl'). "
                             f"Found {scheme}")
        self.mesh = mesh
        self.dt = dt
        self.scheme = scheme
        self.velocity_update = velocity_update
        self.sim_steps = sim_steps
        # solver.py:81-85
        mesh.apply_on_particles("set_particle_element_ids")
        mesh.elements.compute_volume()
        mesh.apply_on_particles("compute_volume", (mesh.elements.total_elements,))

    # scheme.py:22-29
    def compute_nodal_kinematics(self):
        m = self.mesh
        m.apply_on_particles("set_particle_element_ids")
        m.apply_on_particles("update_natural_coords")
        m.apply_on_elements("compute_nodal_mass")
        m.apply_on_elements("compute_nodal_momentum")
        m.apply_on_elements("compute_velocity")
        m.apply_on_elements("apply_boundary_constraints")

    # scheme.py:31-35
    def compute_stress_strain(self):
        m = self.mesh
        m.apply_on_particles("compute_strain", (self.dt,))
        m.apply_on_particles("update_volume")
        m.apply_on_particles("compute_stress")

    # scheme.py:37-56 (t = step * dt in the working dtype, solver.py:213)
    def compute_forces(self, gravity, step):
        m = self.mesh
        dt = m.elements.dtype
        t = dt(step) * dt(self.dt)
        m.apply_on_elements("compute_external_force")
        m.apply_on_elements("compute_body_force", (gravity,))
        m.apply_traction_on_particles(t)
        m.apply_on_elements("apply_concentrated_nodal_forces", (t,))
        m.apply_on_elements("compute_internal_force")

    # scheme.py:58-67
    def compute_particle_kinematics(self):
        m = self.mesh
        m.apply_on_elements("update_nodal_acceleration_velocity", (self.dt,))
        m.apply_on_particles("update_position_velocity",
                             (self.dt, self.velocity_update))

    def step(self, i, gravity):
        """solver.py:179-185; USF stress before forces, USL after G2P
        (scheme.py:98-104, 127-133)."""
        self.compute_nodal_kinematics()
        if self.scheme == "usf":
            self.compute_stress_strain()
        self.compute_forces(gravity, i)
        self.compute_particle_kinematics()
        if self.scheme == "usl":
            self.compute_stress_strain()

    def solve(self, gravity, callback=None, start=0):
        """solver.py:159-221; ``callback(i, self)`` runs after step i (the
        reference's writer hook, solver.py:203-210)."""
        for i in range(start, start + self.sim_steps):
            self.step(i, gravity)
            if callback is not None:
                callback(i, self)
        ps = self.mesh.particles
        return {k: np.stack([getattr(p, k) for p in ps]).squeeze()
                for k in ("loc", "velocity", "stress", "strain")}


# ----------------------------------------------------------------------------
# Convenience builders used by tests and the CPU baseline
# ----------------------------------------------------------------------------
def gauss_points_2d(nx, ny, hx, hy, rows=None, jitter=0.0, seed=0,
                    dtype=np.float32):
    """2x2 Gauss points per cell over the lower ``rows`` cell rows, in cell
    order (row-major) and, within a cell, (-,-), (+,-), (-,+), (+,+).
    Optional uniform jitter of +-jitter*h per coordinate, clamped into the
    mesh interior."""
    rows = ny if rows is None else rows
    g = (1.0 - 1.0 / math.sqrt(3.0)) / 2.0
    offs = np.array([[g, g], [1 - g, g], [g, 1 - g], [1 - g, 1 - g]])
    jj, ii = np.meshgrid(np.arange(rows), np.arange(nx), indexing="ij")
    base = np.stack([ii.reshape(-1), jj.reshape(-1)], axis=1).astype(np.float64)
    pts = (base[:, None, :] + offs[None, :, :]).reshape(-1, 2)
    if jitter:
        rng = np.random.Generator(np.random.PCG64(seed))
        pts = pts + rng.uniform(-jitter, jitter, size=pts.shape)
    pts[:, 0] *= hx
    pts[:, 1] *= hy
    lo = 0.0
    hi_x = np.nextafter(np.float32(nx * hx), np.float32(0))
    hi_y = np.nextafter(np.float32(ny * hy), np.float32(0))
    pts[:, 0] = np.clip(pts[:, 0], lo, hi_x)
    pts[:, 1] = np.clip(pts[:, 1], lo, hi_y)
    return pts.astype(dtype)
