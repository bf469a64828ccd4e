"""CPU oracle for the PDHG hot path -- TEST INFRASTRUCTURE ONLY.

A float64 NumPy/SciPy restatement of the reference's CPU path of
``PDHG(G, F, A).solve(use_gpu=False)`` and of the operators and prox maps it
calls (reference files cited per function).  It is the checker for the HIP
kernels of ``opymize_amd``: only ``tests/``, ``__graft_entry__.smoke()`` and
the ``cpu_baseline`` leg of ``bench.py`` may import it, and never as the
thing measured or shipped.  The product package does not import it.

How each part is pinned (so that a dropped term, a wrong sign or index, or a
transposed operand fails a check that does not just retype the formula):

* ops.grad_forward / div_forward / laplacian / grad_centered: explicit
  per-coordinate loops (``ops_loops``) on small grids, the adjoint identity,
  convergence to analytic derivatives (reference tests/diff.py:42-82) and the
  reference's exact centered-scheme table (tests/sparse.py:59-93).
* ops.matmul / matmul_r: ``numpy.einsum`` with the reference's subscripts.
* prox.l1norms_proj: closed form; the spectral case against SVD clipping.
* prox.quadepi_proj: the geometric invariants of the reference's
  tests/quadepi.py and a brute-force 1-D radial minimisation.
* prox.epigraph_proj: KKT conditions and the constructions of the
  reference's tests/epigraph.py; brute force over active sets.
* pdhg.solve: Chambolle-Pock optimality (fixed point), weak duality
  (gap >= 0), an exact 1-D TV solution by bounded least squares (scipy
  lsq_linear on the dual), a hand-derived two-iteration known answer, and
  the strong-convexity bound 0.5*lambda*|x - x*|^2 <= gap.

* newton (semismooth Newton continuation): the prox Jacobians and the
  Newton matrix against central differences of the prox maps and of the
  fixed-point residual, CG against a direct solve, Armijo against a
  hand-worked backtracking, and descent toward the exact 1-D TV solution.

Parity unpinned: the exact adaptive step trajectory at rounding level
(reference pdhg.py:171,178 branch on FP comparisons; summation order of the
reference's PyCUDA reductions is not fixed by any reference test) -- compared
with tolerances instead.
"""
