"""CPU oracle for the B-spline port-solution path -- TEST INFRASTRUCTURE ONLY.

This package restates, in numpy, what the reference's
``BSplineSolver.compute_port_solution()`` computes (reference
src/momwire/bspline.py:4053-4179 and the helpers it calls).  It exists to check
the HIP path: only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s
``cpu_baseline`` leg may import it, and never as the thing measured or shipped.
The product (``momwire_amd``) must not import anything from here.

How each function is pinned (so that a dropped term, a wrong sign or a
transposed operand fails a check rather than hiding in a retyped formula):

* end-to-end ``y`` / ``coeffs`` / ``Z``: golden vectors written by the
  reference's own solver (tests/golden/make_golden.py), at 1e-10;
* segment-pair moments, same-edge reg moments, static moments and the Galerkin
  assembly: the reference's compiled C++ accelerator built from its sources
  into ``oracle/_ref`` (oracle/Makefile), and brute-force adaptive quadrature
  on tiny inputs;
* static moments additionally: closed-form special cases (a -> 0 limits,
  symmetry J_pq(c) = J_qp(-c) after reflection) and mpmath quadrature;
* basis polynomials: partition of unity and agreement with scipy's B-spline
  evaluator at interior points;
* port algebra: Z X = B residual, KCL rows, reciprocity Y = Y^T.
"""
