"""Seeded synthetic workloads of the benchmark configurations.

``block_2d`` builds the 2D uniform-grid cases: Quadrilateral4Node mesh of
n x n cells on [0,1]^2, 4 particles per cell at the 2x2 Gauss points of the
lower ``rows`` cell rows (optionally jittered uniformly by +-jitter*h per
coordinate, then clamped into [0, X_n) with X_n = fl32(n fl32(1/n)) the last
node coordinate, so that no particle starts outside the mesh),
LinearElastic, gravity (0, -9.81), the bottom node row fixed in x and y and
the left and right node columns fixed in x (rollers).  Without the rollers
the side particles leave the mesh within a few steps under the lateral
expansion of the compressed block; the reference's element search
(element.py:820-840) then assigns them element 0 or wraps them into the next
cell row, and the state diverges.
"""
import math

import numpy as np
import torch


def gauss_points(nx, rows, h, jitter=0.0, seed=0):
    g = (1.0 - 1.0 / math.sqrt(3.0)) / 2.0
    offs = np.array([[g, g], [1 - g, g], [g, 1 - g], [1 - g, 1 - g]], dtype=np.float64)
    jj, ii = np.meshgrid(np.arange(rows, dtype=np.float64), np.arange(nx, dtype=np.float64),
                         indexing="ij")
    base = np.stack([ii.reshape(-1), jj.reshape(-1)], axis=1)
    pts = (base[:, None, :] + offs[None, :, :]).reshape(-1, 2)
    if jitter:
        rng = np.random.Generator(np.random.PCG64(seed))
        pts += rng.uniform(-jitter, jitter, size=pts.shape)
    pts *= h
    # stay below the last node coordinate fl32(n * fl32(h)) (1 for n = 1024):
    # points at or beyond it belong to the next cell row (element.py:820-840)
    last = np.float32(nx) * np.float32(h)
    hi = float(np.nextafter(last, np.float32(0.0)))
    np.clip(pts, 0.0, hi, out=pts)
    return pts


def gaussian_field(n, h, corr, seed):
    """Unit-variance Gaussian random field on the (n+1)^2 node grid (indexed
    [j, i]): seeded white noise filtered by a Gaussian of standard deviation
    corr (in length units, periodic), normalised to zero mean and unit
    variance."""
    from scipy.ndimage import gaussian_filter
    rng = np.random.Generator(np.random.PCG64(seed))
    w = rng.standard_normal((n + 1, n + 1))
    f = gaussian_filter(w, sigma=corr / h, mode="wrap")
    return (f - f.mean()) / f.std()


def bilinear(field, h, pts):
    """Bilinear interpolation of a node-grid field (indexed [j, i]) at points."""
    x = pts[:, 0] / h
    y = pts[:, 1] / h
    i = np.clip(np.floor(x).astype(int), 0, field.shape[1] - 2)
    j = np.clip(np.floor(y).astype(int), 0, field.shape[0] - 2)
    fx, fy = x - i, y - j
    return ((1 - fx) * (1 - fy) * field[j, i] + fx * (1 - fy) * field[j, i + 1]
            + (1 - fx) * fy * field[j + 1, i] + fx * fy * field[j + 1, i + 1])


def inverse_2d(n=128, seed=0, dtype=torch.float32, device="cuda", side_rollers=False):
    """The differentiable inverse problem's case: block_2d(n) with particles
    at the 2x2 Gauss points of the lower n/2 cell rows (no jitter), nu = 0.3,
    rho = 1000, gravity, bottom nodes fixed, dt = 1e-5, USF, and the true
    per-particle modulus E_p = 1e6 exp(0.3 g(x_p)), g = gaussian_field(n, 1/n,
    0.1, seed) interpolated bilinearly.  The simulation starts with the
    uniform E = 1e6.  Returns (MPMExplicit, gravity, E_true as an (N,) float64
    CPU tensor)."""
    sim, g = block_2d(n, rows=n // 2, jitter=0.0, E=1e6, nu=0.3, scheme="usf", dtype=dtype,
                      device=device, side_rollers=side_rollers)
    h = 1.0 / n
    pts = gauss_points(n, n // 2, h)
    E = 1e6 * np.exp(0.3 * bilinear(gaussian_field(n, h, 0.1, seed), h, pts))
    return sim, g, torch.from_numpy(E)


def ensemble_specs(nsims=64, seed=0):
    """Member parameters of the uneven ensemble: n = round(exp(U(ln 32,
    ln 256))) cells per side, rows = ceil(U(0.3, 1.0) n) particle rows,
    E = exp(U(ln 1e5, ln 1e7)), nu = U(0.2, 0.4); one seeded generator."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for _ in range(nsims):
        n = int(round(math.exp(rng.uniform(math.log(32), math.log(256)))))
        rows = int(math.ceil(rng.uniform(0.3, 1.0) * n))
        E = float(math.exp(rng.uniform(math.log(1e5), math.log(1e7))))
        nu = float(rng.uniform(0.2, 0.4))
        out.append(dict(n=n, rows=min(rows, n), E=E, nu=nu))
    return out


def ensemble(nsims=64, seed=0, scheme="usl", dtype=torch.float32, device="cuda"):
    """The uneven ensemble: ``nsims`` block_2d members (particles at the 2x2
    Gauss points of the cells of the lower ``rows`` rows, no jitter, bottom
    fixed, side rollers, gravity, rho = 1000, dt = 1e-5, USL).  Returns (list
    of MPMExplicit, gravity, member parameters)."""
    specs = ensemble_specs(nsims, seed)
    sims = []
    for k, m in enumerate(specs):
        sim, g = block_2d(m["n"], rows=m["rows"], jitter=0.0, seed=seed + 1 + k, E=m["E"],
                          nu=m["nu"], scheme=scheme, dtype=dtype, device=device)
        sims.append(sim)
    return sims, g, specs


def block_2d(n=1024, rows=None, jitter=0.25, seed=0, E=1e6, nu=0.3, density=1000.0,
             dt=1e-5, scheme="usf", velocity_update=False, dtype=torch.float32,
             device="cuda", keep_temporaries=False, fix_bottom=(0, 1), side_rollers=True,
             round_positions=None):
    """Returns (MPMExplicit, gravity) for the uniform 2D block case.  The
    4.2M-particle throughput configuration is ``block_2d(1024)``.
    round_positions (a NumPy dtype): round the generated positions to it
    first, so that blocks built in different precisions start from the same
    positions, cells and particle volumes."""
    import diffmpm_amd as d
    rows = n if rows is None else rows
    h = 1.0 / n
    pts = gauss_points(n, rows, h, jitter, seed)
    if round_positions is not None:
        pts = pts.astype(round_positions).astype(np.float64)
    bottom = torch.arange(n + 1)
    cons = [(bottom, d.Constraint(k, 0.0)) for k in fix_bottom]
    if side_rollers:
        j = torch.arange(n + 1)
        cons.append((torch.cat([j * (n + 1), j * (n + 1) + n]), d.Constraint(0, 0.0)))
    el = d.Quadrilateral4Node((n, n), n * n, (h, h), cons, device=device, dtype=dtype)
    mat = d.LinearElastic({"density": density, "youngs_modulus": E, "poisson_ratio": nu})
    np_dt = np.float64 if dtype == torch.float64 else np.float32
    loc = torch.from_numpy(pts.astype(np_dt)).reshape(-1, 1, 2)
    ps = d.Particles(loc, mat, torch.zeros(loc.shape[0], dtype=torch.int32), device=device,
                     dtype=dtype)
    mesh = d.Mesh2D({"particles": [ps], "elements": el, "particle_surface_traction": []})
    sim = d.MPMExplicit(mesh, dt, scheme, velocity_update, 1, keep_temporaries=keep_temporaries)
    return sim, torch.tensor([0.0, -9.81], dtype=torch.float64)
