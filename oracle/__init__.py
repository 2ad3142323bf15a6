"""CPU oracle for the equiprop path: test infrastructure, never part of the product.

Importable only from tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg.
"""
