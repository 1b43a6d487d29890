"""CPU oracle for the dycore step -- TEST INFRASTRUCTURE ONLY.

A NumPy restatement (float32 or float64) of the reference's
``gpuwm.core.dycore.step`` and the kernels it launches, used as the checker
for the HIP implementation in ``arwen_amd``.  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it; the product path never does.

Pinning (what makes the restatement trustworthy, beyond its own text):

* advection: ``tests/golden/advection_periodic_regression.npz`` (inputs and
  outputs of the reference's four flux-divergence kernels);
* whole step: ``tests/golden/phase2_step_regression.npz`` (``dry_flat`` and
  ``terrain`` states after 5 reference ``dycore.step`` calls, rebuilt from
  the same initial-state builders);
* closed forms and invariants: implicit-w tridiagonal vs a dense solve,
  discrete hydrostatic balance (rest state stays at rest), periodic mass
  telescoping, FP32-vs-FP64 agreement, Straka density-current gates, sound
  speed.

Each function cites the reference file:line it follows.
"""
