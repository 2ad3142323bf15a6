"""Tests shipped with the package (``pytest --pyargs parament_amd``), covering what the reference
package's own test modules check (pyparament/parament/test: test_error.py, test_numerics.py,
test_wrapper.py).  Tests marked ``gpu`` need a HIP device; ``-m "not gpu"`` runs the rest.
The repository's ``tests/`` directory holds the full suite (oracle, parity, gradient, ensembles)."""
