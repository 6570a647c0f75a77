"""The oracle's 2D Quadrilateral4Node step restated with torch CPU tensors —
TEST INFRASTRUCTURE ONLY (the multi-threaded CPU baseline of ``bench.py``).

It computes what ``oracle.mpm.MPMExplicit.step`` computes for one particle
set on a Quadrilateral4Node mesh with LinearElastic particles, no particle
tractions and no concentrated nodal forces (the throughput workload): the
same phases in the same order (scheme.py:78-133), the same formulas, with
torch ops that run on every host thread.  Scatters use ``index_add_``, so
their accumulation order is not the reference's particle order; the results
agree with the NumPy oracle to FP32 rounding (``tests/test_oracle_torch.py``
checks it on seeded cases, USF and USL).  Reference lines are cited per
phase.
"""
import torch

SNAP = 1e-12  # element.py:186,199,321,326; particle.py:311


def _snap(x):
    return torch.where(x.abs() < SNAP, torch.zeros_like(x), x)


class TorchQ4:
    """One particle set on an nx x ny Q4 mesh (node n = j (nx+1) + i at
    (fl(i hx), fl(j hy)), element.py:667-723).  constraints: list of (node
    ids, dir, velocity) applied in order (constraint.py:29-44).  de: the 6x6
    De of material.py:94-111 in double."""

    def __init__(self, nx, ny, hx, hy, loc, de, density, constraints, dt, scheme="usf",
                 velocity_update=False, dtype=torch.float32):
        self.nx, self.ny, self.dtype = nx, ny, dtype
        self.hx, self.hy = torch.tensor(hx, dtype=dtype), torch.tensor(hy, dtype=dtype)
        self.X = torch.arange(nx + 1).to(dtype) * self.hx
        self.Y = torch.arange(ny + 1).to(dtype) * self.hy
        jj, ii = torch.meshgrid(torch.arange(ny + 1), torch.arange(nx + 1), indexing="ij")
        self.node_loc = torch.stack([self.X[ii.reshape(-1)], self.Y[jj.reshape(-1)]], 1)
        M = self.node_loc.shape[0]
        self.M = M
        self.cons = [(torch.as_tensor(i).long(), int(d), float(v)) for i, d, v in constraints]
        self.de = torch.as_tensor(de, dtype=torch.float64).to(dtype)
        self.dt = torch.tensor(dt, dtype=dtype)
        self.scheme, self.vu = scheme, bool(velocity_update)
        n = loc.shape[0]
        self.loc = torch.as_tensor(loc).to(dtype).clone()
        z2 = torch.zeros((n, 2), dtype=dtype)
        self.velocity = z2.clone()
        self.stress = torch.zeros((n, 6), dtype=dtype)
        self.strain = torch.zeros((n, 6), dtype=dtype)
        self.evc = torch.zeros(n, dtype=dtype)
        self.density = torch.full((n,), float(density), dtype=dtype)
        self.nv = torch.zeros((M, 2), dtype=dtype)   # nodal velocity (kept between steps)
        # solver.py:81-85: element ids, element volume, V_p = V_e / count(e),
        # m = V rho (particle.py:158-178)
        self.eid = self.element_ids(self.loc)
        a = c = self.hy
        b = d = self.hx
        p = q = torch.sqrt(a * a + b * b)
        ve = torch.tensor(0.25, dtype=dtype) * torch.sqrt(
            torch.tensor(4, dtype=dtype) * p * p * q * q - (a * a + c * c - b * b - d * d) ** 2)
        cnt = torch.bincount(self.eid.clamp(min=0, max=nx * ny - 1).long(),
                             minlength=nx * ny).to(dtype)
        self.volume = ve / cnt[self.eid.clamp(0, nx * ny - 1).long()]
        self.mass = self.volume * self.density

    # element.py:820-840 (closed form, see oracle.mpm.Elements.element_ids)
    def element_ids(self, loc):
        x, y = loc[:, 0].contiguous(), loc[:, 1].contiguous()
        I = torch.searchsorted(self.X, x, right=True) - 1
        J = torch.searchsorted(self.Y, y, right=True) - 1
        bad = (I < 0) | (J < 0) | torch.isnan(x) | torch.isnan(y)
        return torch.where(bad, torch.zeros_like(I), J * self.nx + I)

    def _ids(self):
        e = self.eid.long()
        nx = self.nx
        ll = (e // nx) * (nx + 1) + e % nx
        return torch.stack([ll, ll + 1, ll + nx + 2, ll + nx + 1], 1)

    def _gather(self, arr, ids):
        return arr[ids.clamp(0, arr.shape[0] - 1)]

    def _scatter(self, ids, vals):
        """arr.at[ids].add(vals) (JAX drops out-of-range ids) into a fresh
        (M, k) array."""
        out = torch.zeros((self.M, vals.shape[-1]), dtype=self.dtype)
        i = ids.reshape(-1)
        v = vals.reshape(-1, vals.shape[-1])
        ok = (i >= 0) & (i < self.M)
        out.index_add_(0, i[ok], v[ok])
        return out

    @staticmethod
    def _N(xi):
        x, y = xi[:, 0], xi[:, 1]
        q = 0.25
        return torch.stack([q * (1 - x) * (1 - y), q * (1 + x) * (1 - y),
                            q * (1 + x) * (1 + y), q * (1 - x) * (1 + y)], 1)

    @staticmethod
    def _dNdxi(xi):
        x, y = xi[:, 0], xi[:, 1]
        g = torch.empty((xi.shape[0], 4, 2), dtype=xi.dtype)
        g[:, 0, 0] = -0.25 * (1 - y); g[:, 0, 1] = -0.25 * (1 - x)
        g[:, 1, 0] = 0.25 * (1 - y);  g[:, 1, 1] = -0.25 * (1 + x)
        g[:, 2, 0] = 0.25 * (1 + y);  g[:, 2, 1] = 0.25 * (1 + x)
        g[:, 3, 0] = -0.25 * (1 + y); g[:, 3, 1] = 0.25 * (1 - x)
        return g

    def _dNdx(self, xi, coords):
        """element.py:791-818: J = dN/dxi^T X_e, dN/dx = dN/dxi inv(J)^T."""
        g = self._dNdxi(xi)
        J = torch.einsum("nka,nkb->nab", g, coords)
        a, b, c, d = J[:, 0, 0], J[:, 0, 1], J[:, 1, 0], J[:, 1, 1]
        det = a * d - b * c
        inv = torch.stack([torch.stack([d, -b], 1), torch.stack([-c, a], 1)], 1) / det[:, None, None]
        return torch.einsum("nka,nca->nkc", g, inv)

    def _rate(self, dN, ve):
        """particle.py:281-313: L = dN^T v_e, [L00, L11, 0, L01 + L10, 0, 0]."""
        L = torch.einsum("nka,nkb->nab", dN, ve)
        r = torch.zeros((dN.shape[0], 6), dtype=self.dtype)
        r[:, 0] = L[:, 0, 0]
        r[:, 1] = L[:, 1, 1]
        r[:, 3] = L[:, 0, 1] + L[:, 1, 0]
        return _snap(r)

    def _constrain(self, v, p=None, a=None, m=None):
        for ids, d, val in self.cons:  # constraint.py:29-44
            v[ids, d] = val
            if p is not None:
                p[ids, d] = m[ids] * torch.tensor(val, dtype=self.dtype)
            if a is not None:
                a[ids, d] = 0

    def _stress_strain(self, ids, coords, xi):
        """particle.py:249-327, material.py:113-116."""
        ve = self._gather(self.nv, ids)
        rate = self._rate(self._dNdx(xi, coords), ve)
        dstrain = rate * self.dt
        self.strain = self.strain + dstrain
        rc = self._rate(self._dNdx(torch.zeros_like(xi), coords), ve)
        dvol = self.dt * (rc[:, 0] + rc[:, 1])
        self.evc = self.evc + dvol
        f = 1 + dvol
        self.volume = self.volume * f
        self.density = self.density / f
        self.stress = self.stress + dstrain @ self.de.T

    def step(self, i, gravity):
        dt = self.dt
        # scheme.py:22-29: element ids, natural coords, nodal m, p, v, constraints
        self.eid = self.element_ids(self.loc)
        ids = self._ids()
        coords = self._gather(self.node_loc, ids)
        t0, t2 = coords[:, 0], coords[:, 2]
        xi = (self.loc - (t0 + t2) / 2) * (2 / (t2 - t0))     # particle.py:180-206
        N = self._N(xi)
        nm = self._scatter(ids, (self.mass[:, None] * N)[..., None])[:, 0]
        pm = self.mass[:, None] * self.velocity
        mom = _snap(self._scatter(ids, N[:, :, None] * pm[:, None, :]))
        m = nm[:, None]
        v = _snap(torch.where(m == 0, self.nv, mom / m))      # element.py:191-202
        self._constrain(v, mom, None, nm)
        self.nv = v
        if self.scheme == "usf":                              # scheme.py:98-104
            self._stress_strain(ids, coords, xi)
        # scheme.py:37-56: external (particle f_ext = 0), body and internal forces
        zero = torch.zeros((len(self.loc), 2), dtype=self.dtype)
        f_ext = self._scatter(ids, N[:, :, None] * zero[:, None, :])
        g = torch.as_tensor(gravity, dtype=self.dtype)
        f_ext = f_ext + self._scatter(ids, N[:, :, None] * (self.mass[:, None] * g)[:, None, :])
        dN = self._dNdx(xi, coords)
        s = self.stress
        fx = dN[:, :, 0] * s[:, 0:1] + dN[:, :, 1] * s[:, 3:4]
        fy = dN[:, :, 1] * s[:, 1:2] + dN[:, :, 0] * s[:, 3:4]
        nV = -self.volume[:, None]
        f_int = self._scatter(ids, torch.stack([nV * fx, nV * fy], 2))
        # element.py:305-329: nodal acceleration and velocity, constraints
        a = torch.nan_to_num((f_int + f_ext) / m)
        v = self.nv + a * dt
        self._constrain(v, None, a, nm)
        self.nv = _snap(v)
        a = _snap(a)
        # particle.py:208-247: G2P
        nvp = (N[:, :, None] * self._gather(self.nv, ids)).sum(1)
        nap = (N[:, :, None] * self._gather(a, ids)).sum(1)
        self.velocity = nvp if self.vu else self.velocity + nap * dt
        self.loc = self.loc + nvp * dt
        if self.scheme == "usl":                              # scheme.py:127-133
            self._stress_strain(ids, coords, xi)
