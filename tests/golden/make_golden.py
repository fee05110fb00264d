"""Write the end-to-end golden vectors for the port-solution path.

Runs ONLY in the development container, where the reference checkout is
mounted read-only at /root/reference.  It imports the reference's pure-Python
solver (its numpy fallback: no compiled accelerator is loaded) and records,
for the structures in ``momwire_amd.workloads``, what
``BSplineSolver.compute_port_solution()`` returns (reference
src/momwire/bspline.py:4053-4179): the port admittance matrix ``y`` and the
coefficient columns ``coeffs``; for the smallest case also the dense Z.

Nothing else pins Y and coeffs end to end at 1e-10, so these vectors are what
the oracle and the GPU path are both held to.  Outputs are small ``.npz``
files next to this script.

    python tests/golden/make_golden.py [name ...]
"""

import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
sys.path.insert(0, "/root/reference/src")

import momwire  # noqa: E402
from momwire import BSplineSolver  # noqa: E402

from momwire_amd import workloads  # noqa: E402


def _solve(kw):
    s = BSplineSolver(**kw)
    t0 = time.time()
    sol = s.compute_port_solution()
    return s, sol, time.time() - t0


def case_c1():
    kw = workloads.straight_dipole()
    s, sol, dt = _solve(kw)
    geom = s._build_geometry()
    supp, polys, kcl, _wk, _wb = s._build_basis_polynomials(geom)
    Z = s._compute_Z_operator(geom, supp, polys)
    return dict(y=sol.y, coeffs=sol.coeffs, Z=np.asarray(Z), k=s.k,
                omega=s.omega), dt


def case_c2():
    kw = workloads.dipole_array(8, 8, 41, seed=1)
    s, sol, dt = _solve(kw)
    cols = np.array([0, 9, 27, 63])
    return dict(y=sol.y, coeffs_cols=cols, coeffs=sol.coeffs[:, cols],
                coeffs_colsum=sol.coeffs.sum(axis=1), k=s.k), dt


def case_small_array():
    kw = workloads.dipole_array(2, 2, 11, seed=5)
    s, sol, dt = _solve(kw)
    return dict(y=sol.y, coeffs=sol.coeffs), dt


def _tee(ports):
    kw = workloads.junction_tee(ports)
    s, sol, dt = _solve(kw)
    out = dict(y=sol.y, coeffs=sol.coeffs)
    if ports == "feed":
        ks = np.array([2 * np.pi * 0.9, 2 * np.pi * 1.0, 2 * np.pi * 1.15])
        out["sweep_k"] = ks
        out["sweep_y"] = s.compute_y_matrix_swept(ks)
        geom = s._build_geometry()
        supp, polys, kcl, _wk, _wb = s._build_basis_polynomials(geom)
        out["Z"] = np.asarray(s._compute_Z_operator(geom, supp, polys))
        out["knot_currents"] = np.concatenate(s.currents_at_knots(sol.coeffs[:, 0]))
    return out, dt


def case_sweep_struct(index):
    """One random structure of the batched sweep, swept over 3 k."""
    kw = workloads.random_structure(3 * 100003 + index)
    s = BSplineSolver(**kw)
    ks = workloads.sweep_wavenumbers()[[0, 7, 15]]
    t0 = time.time()
    ys = s.compute_y_matrix_swept(ks)
    sol = s.compute_port_solution()
    return dict(sweep_k=ks, sweep_y=ys, y=sol.y, coeffs=sol.coeffs), time.time() - t0


CASES = {
    "sweep_struct4": lambda: case_sweep_struct(4),
    "sweep_struct8": lambda: case_sweep_struct(8),
    "c1_dipole": case_c1,
    "c2_array8x8": case_c2,
    "array2x2": case_small_array,
    "tee_feed": lambda: _tee("feed"),
    "tee_jport": lambda: _tee("junction_port"),
    "tee_nodegap": lambda: _tee("node_gap"),
}


def main(names):
    assert not momwire.accelerated, "expected the reference's pure-Python path"
    for name in names or CASES:
        out, dt = CASES[name]()
        path = os.path.join(HERE, f"{name}.npz")
        np.savez_compressed(path, **out)
        print(f"{name}: {dt:.1f} s -> {os.path.relpath(path)}", flush=True)


if __name__ == "__main__":
    main(sys.argv[1:])
