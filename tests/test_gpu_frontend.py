"""The reference's 2D benchmark input files (TOML, entity sets, particle
JSON, copied unmodified into tests/golden/benchmarks/) run through the TOML
front-end (``diffmpm_amd.MPM``, the reference's ``diffmpm.MPM``) on the GPU
path, and the NPZ files it writes meet the reference's own assertions
(benchmarks/2d/*/test_benchmark.py, restated in reference_goldens.json)."""
import os
import shutil

import numpy as np
import pytest

import cases

pytestmark = pytest.mark.gpu

GOLDENS = cases.load_goldens()
FILES = {"uniaxial_stress": "mpm-uniaxial-stress.toml",
         "uniaxial_nodal_forces": "mpm-nodal-forces.toml",
         "uniaxial_particle_traction": "mpm-particle-traction.toml"}


def _npz_name(step, nsteps):
    digits = int(np.log10(nsteps)) + 1
    return f"particles_{step:0{digits}d}.npz"


@pytest.mark.parametrize("name", sorted(FILES))
def test_benchmark_input_files_end_to_end(name, tmp_path, monkeypatch):
    src = os.path.join(cases.GOLDEN_DIR, "benchmarks", name)
    for f in os.listdir(src):
        shutil.copyfile(os.path.join(src, f), tmp_path / f)
    monkeypatch.chdir(tmp_path)  # as the reference's benchmark tests do
    from diffmpm_amd import MPM
    mpm = MPM(FILES[name])
    out = mpm.solve()
    assert np.isfinite(out["loc"].cpu().numpy()).all()
    meta = mpm._config.parsed_config["meta"]
    folder = tmp_path / "results" / meta["title"]
    for chk in GOLDENS[name]["checks"]:
        with np.load(folder / _npz_name(chk["after_step"], meta["nsteps"])) as r:
            assert set(r.files) == {"loc", "velocity", "stress", "strain"}
            got, ok = cases.check_value(r["stress"][0], chk)
        assert ok, (chk, got)


def test_config_parses_sets_functions_and_loads(tmp_path, monkeypatch):
    """Parsed structures of the traction benchmark: node and particle sets,
    the Linear function, the traction and constraint records."""
    src = os.path.join(cases.GOLDEN_DIR, "benchmarks", "uniaxial_particle_traction")
    from diffmpm_amd.io import Config
    cfg = Config(os.path.join(src, FILES["uniaxial_particle_traction"]))  # files found beside it
    pc = cfg.parsed_config
    assert pc["meta"]["nsteps"] == 1000 and pc["meta"]["scheme"] == "usf"
    (pt,) = pc["particle_surface_traction"]
    assert pt.pids.tolist() == [5, 11] and pt.dir == 0 and pt.traction == 1
    assert list(pt.function.xvalues) == [0.0, 0.5, 1.0]
    el = pc["elements"]
    assert el.nelements == (3, 1)
    (ids, c), = el.constraints
    assert ids.tolist() == [0, 4] and c.dir == 0 and c.velocity == 0.0
    (ps,) = pc["particles"]
    assert len(ps) == 12 and float(ps.velocity.abs().max()) == 0.0
