"""The benchmark configurations as oracle cases (CPU): their specs build, and
the 1D bar is CFL-unstable as its configuration notes (c dt / h = 12.8)."""
import math

import numpy as np

import cases


def test_c1_bar_oracle_turns_non_finite():
    spec = cases.config_c1_spec(nsteps=200)
    osim = cases.build_oracle(spec, np.float64)
    g = np.asarray(spec["gravity"])
    first_bad = None
    for i in range(spec["nsteps"]):
        osim.step(i, g)
        if not np.all(np.isfinite(osim.mesh.particles[0].loc)):
            first_bad = i
            break
    assert first_bad is not None and first_bad < 200, first_bad
    c = math.sqrt(100.0 / 1.0)
    assert c * spec["dt"] / spec["el_len"] > 10


def test_c4_youngs_field():
    """The random field is unit-variance and zero-mean on the node grid,
    spatially correlated over ~0.1 (neighbouring nodes nearly equal), and the
    bilinear interpolation reproduces it at the nodes."""
    n, h = 128, 1.0 / 128
    f = cases.gaussian_random_field(n, h, 0.1, 0)
    assert abs(f.mean()) < 1e-12 and abs(f.std() - 1.0) < 1e-12
    assert np.corrcoef(f[:, :-1].ravel(), f[:, 1:].ravel())[0, 1] > 0.99
    jj, ii = np.meshgrid(np.arange(0, n, 7), np.arange(0, n, 5), indexing="ij")
    pts = np.stack([ii.ravel() * h, jj.ravel() * h], axis=1)
    assert np.allclose(cases.bilinear(f, h, pts), f[jj.ravel(), ii.ravel()], atol=1e-12)
    spec = cases.config_c4_spec()
    E = np.asarray(spec["youngs"])
    assert len(E) == 128 * 64 * 4
    assert np.allclose(E, 1e6 * np.exp(0.3 * cases.bilinear(f, h, np.asarray(spec["particles"]))))


def test_npz_file_names_follow_the_reference():
    """writers.npz_name: particles_<step>.npz with the step zero-padded to the
    number of digits of max_steps (diffmpm/writers.py:45-52)."""
    import math
    from diffmpm_amd.writers import npz_name
    for max_steps in (1, 9, 10, 99, 100, 999, 1000, 12345):
        digits = int(math.log10(max_steps)) + 1
        for step in sorted({0, 1, 7, max_steps // 2, max_steps - 1}):
            want = "particles_" + str(step).rjust(digits, "0") + ".npz"
            assert npz_name(step, max_steps) == want, (step, max_steps)
