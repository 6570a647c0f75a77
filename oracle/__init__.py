"""CPU restatement of diffmpm's explicit MPM step — TEST INFRASTRUCTURE ONLY.

This package is the oracle that the HIP path of ``diffmpm_amd`` is checked
against.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import it, and only as the source of
the values the HIP path is compared with (or, for ``cpu_baseline``, as the
timed CPU column).  The product package
``diffmpm_amd`` never imports it and has no CPU fallback.

``oracle.mpm`` restates, in NumPy, the arithmetic of the reference's
``diffmpm/element.py``, ``particle.py``, ``node.py``, ``material.py``,
``constraint.py``, ``functions.py``, ``mesh.py``, ``scheme.py`` and
``solver.py``; each function cites the reference lines it follows.  It is
templated on dtype: float32 mirrors the reference (JAX with x64 disabled),
float64 is the high-precision oracle.  P2G scatters use ``np.add.at``, which
accumulates in particle-index order like the reference's serial
``lax.fori_loop`` scatter loops.

Pinning (how each function is checked, see ``tests/test_oracle_*.py``):

* every unit-test known-answer value of the reference
  (``tests/test_element.py``, ``tests/test_material.py``,
  ``tests/test_particle.py``) restated as fixtures in
  ``tests/golden/reference_kats.json``;
* the three end-to-end goldens of ``benchmarks/2d/*/test_benchmark.py``
  restated in ``tests/golden/reference_goldens.json``;
* closed forms and invariants: partition of unity, gradient sums, P2G mass and
  momentum conservation, exact strain rates for linear velocity fields,
  De algebra, ``np.interp`` semantics, brute-force element search on tiny
  meshes, FP32 vs FP64 agreement;
* linear completeness, which fixes the natural coordinates and G2P of both
  elements (``TestLinear1D`` in the reference is empty): sum_i X_i N_i(xi_p)
  = x_p, G2P of a linear nodal velocity gives a x_p exactly, the P2G first
  moments sum_i X_i m_i = sum_p m_p x_p, and the internal-force virial; a
  planted mistake in the 1D natural coordinate (negated, scaled, centred on
  node 0) fails them;
* the per-particle E of SimpleMaterial (an extension): equal to the scalar
  material with that E.

Parity unpinned (no reference test covers them and the reference cannot be
imported here — JAX is absent):

* element search for points outside the mesh (wrap / element-0 / out-of-range
  ids), checked by brute force against the reference's set definition only;
* the exact FP32 rounding of XLA fusions (FMA contraction, LU inverse), pinned
  to tolerance only.
"""
