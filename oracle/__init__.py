"""CPU oracle for the swept / standard stencil path.

TEST INFRASTRUCTURE ONLY.  Nothing in ``pysweep_amd`` imports this package;
only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` use it, and only as the checker.

Contents (each function cites the reference file:line it restates):

* ``heat``    -- heat-equation step in the operation order of the reference's
                 numpy plugin (pysweep/equations/heat.py:48-64) and of its
                 device plugin (pysweep/equations/heat.cu:11-48); initial
                 condition, analytic decay and the exact discrete Fourier
                 solution of the linear scheme.
* ``euler``   -- 2-D Euler finite-volume step in the operation order of
                 pysweep/equations/euler.py:131-313 ("py" order) and of
                 pysweep/equations/euler.cu:22-244 ("cu" order); isentropic
                 vortex initial condition / exact solution (euler.py:63-116).
* ``plugins`` -- the known-answer plugins example / checker / half
                 (pysweep/equations/example.*, checker.*, half.*).
* ``swept``   -- the swept space-time schedule (block.py:123-164,
                 functions.py:12-123, solver.py:113-176, pysweep.cu:64-249)
                 and the standard stage loop (solver.py:178-189), on a
                 periodic grid, with a read-after-write checker.

How each function is pinned (no function here is "parity unpinned"):

* heat py/cu order -- exact discrete Fourier solution of the linear scheme
  (to ~1e-13), the analytic decay exp(-8 pi^2 alpha t) (O(dx^2)), and the
  golden vectors written from the reference's own heat.py
  (tests/golden/make_golden.py).
* euler py order   -- golden vectors written from the reference's own
  euler.py (tests/golden/make_golden.py), free-stream preservation,
  conservation of the four integrals, x<->y mirror symmetry, and convergence
  to the advected vortex.
* euler cu order   -- agreement with the py order to round-off per stage, the
  same invariants, and a planted-mistake check in tests.
* swept schedule   -- swept == standard bitwise for every plugin, the
  reference's known-answer tests (tests.py:255-408) and a read-after-write
  check that every (cell, level) is written exactly once before it is read.
"""
