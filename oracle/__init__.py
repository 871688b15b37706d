"""CPU FP64 restatement of the reference's SIMPLE hot path -- TEST INFRASTRUCTURE.

This package is the oracle.  Only ``tests/``, ``__graft_entry__.smoke()`` and
``bench.py``'s ``cpu_baseline`` leg may import it, and only as the checker or
the reported CPU baseline -- never as a product code path.  The product
(``karman_taichi_amd``) runs hand-written HIP kernels and raises if they are
missing.

Contents
--------
``fvm``     NumPy restatement of the discretisation: u/v/p' coefficient
            assembly with boundary conditions and obstacle mask, mass
            imbalance, corrections, momentum blend, outlet flux fix, time
            advance and the SIMPLE outer step (simple.py, karman_taichi.py).
``krylov``  NumPy restatement of the 5-point stencil, the reductions, BiCGSTAB
            (bicgsolver.py) and CG (cgsolver.py), plus a flexible
            right-preconditioned BiCGSTAB and CSR export for scipy checks.
``cpu_simple.c``  C/OpenMP FP64 restatement with the reference's unfused
            per-kernel decomposition, timed as the CPU baseline.

How each function is pinned (closed forms, scipy, invariants, brute force,
Ghia 1982) is listed in the tests that exercise it.  Taichi itself
(``taichi==0.9.0``, requirements.txt:1) is not installed and cannot be, so the
last-bit behaviour of the reference (fast-math, reduction order, the
out-of-bounds halo reads of cgsolver.py:59-60) is parity unpinned.
"""
