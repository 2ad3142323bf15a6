"""Writes tests/golden/equiprop_golden.npz: known-answer vectors for the equiprop path.

Contents (all complex128 unless noted):
* ``doc_*``        -- the docstring example of the reference Parament class
                      (parament.py:61-68): H0 = diag(1,-1), H1 = sigma_x, dt = 1, c = 0,
                      expected diag(0.54030234-0.84147096j, 0.54030234+0.84147096j)
                      as printed there (8 digits).
* ``seed27_d{2,16}_*`` -- the reference's random_hamiltonians(dim) draw
                      (test_numerics.py:24-29, legacy np.random.seed(27) stream) and
                      scipy.linalg.expm(-1j * 0.01 * H0), its single-step expected value
                      (test_numerics.py:96, 137).
* ``multi_d{2,16}_*`` -- the 10-step two-field case (test_numerics.py:100-117): the two
                      control draws np.random.rand(10) that follow the seed-27 draw, and the
                      stepwise scipy product the reference compares against.
* ``mode_*``       -- oracle outputs for every quadrature mode (d = 4, pts = 9, two
                      controls, seeded) as regression vectors for the restatement.

Run from the repository root: ``python tests/golden/make_golden.py``.
"""
import os
import sys

import numpy as np
import scipy.linalg

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from oracle import equiprop_oracle as eo  # noqa: E402


def random_hamiltonians(dim):
    np.random.seed(27)
    h0 = np.random.uniform(-1, 1, (dim, dim)) / dim + 1j * np.random.uniform(-1, 1, (dim, dim)) / 2
    h0 = h0 + h0.conj().T
    return h0, np.zeros((dim, dim), dtype=complex)


def main():
    out = {}
    out["doc_h0"] = np.diag([1.0, -1.0]).astype(complex)
    out["doc_h1"] = np.array([[0, 1], [1, 0]], dtype=complex)
    out["doc_expected"] = np.diag([0.54030234 - 0.84147096j, 0.54030234 + 0.84147096j])
    for dim in (2, 4, 16):
        h0, _ = random_hamiltonians(dim)
        out[f"seed27_d{dim}_h0"] = h0
        out[f"seed27_d{dim}_expm"] = scipy.linalg.expm(-1j * 0.01 * h0)
        out[f"seed27_d{dim}_expm_debug"] = scipy.linalg.expm(h0)
    for dim in (2, 16):
        h0, h1 = random_hamiltonians(dim)
        c1 = np.random.rand(10)
        c2 = np.random.rand(10)
        u = np.eye(dim, dtype=complex)
        for i in range(10):
            u = scipy.linalg.expm(-1j * (h0 + c1[i] * h1 + c2[i] * h0) * 0.01) @ u
        out[f"multi_d{dim}_c"] = np.stack([c1, c2])
        out[f"multi_d{dim}_expected"] = u
    rng = np.random.default_rng(7)
    d = 4
    g = rng.normal(size=(3, d, d)) + 1j * rng.normal(size=(3, d, d))
    hs = (g + np.conj(np.transpose(g, (0, 2, 1)))) / 4
    c = rng.uniform(-1, 1, (2, 9))
    out["mode_h"] = hs
    out["mode_c"] = c
    for name, mode, mag in (("none", "none", False), ("midpoint", "midpoint", False),
                            ("simpson", "simpson", False), ("magnus", "simpson", True)):
        for prec in ("fp32", "fp64"):
            out[f"mode_{name}_{prec}"] = eo.equiprop(hs[0], list(hs[1:]), c, 0.05, mode=mode, use_magnus=mag,
                                                     precision=prec)
    np.savez_compressed(os.path.join(os.path.dirname(os.path.abspath(__file__)), "equiprop_golden.npz"), **out)


if __name__ == "__main__":
    main()
