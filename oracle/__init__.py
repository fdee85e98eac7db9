"""CPU oracle for the YASPS Newton linear-solve path.

TEST INFRASTRUCTURE ONLY. Nothing in the product package imports this package:
only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` use it, and only as the checker (or the timed CPU baseline), never as
a fallback for the HIP path.

Every function restates the reference algorithm it names (file:line relative to
the reference repository root) in plain numpy. How each part is pinned:

* ``energies``   - SymPy golden vectors (the 6-tet cube fixture of
                   ``tests/energy/check_stable_neo_hookean.py:60-104``), central finite
                   differences, rigid-motion invariants.
* ``projection`` - ``numpy.linalg.eigh`` clamp rules, brute-force PSD checks, the
                   N==1 special case, the Jacobi pseudo-inverse rule.
* ``indices``    - dense brute-force scatter of tiny meshes (as the reference test
                   ``tests/differentiator/test_hessian_sparsity_permutation_gpu.py:104-106``
                   does), hand-counted layouts for repeated vertices and union padding.
* ``assembly``   - dense ``np.add.at`` reference assembly; symmetric SpMV against a
                   dense product.
* ``pcg``        - ``numpy.linalg.solve`` / residual checks; iteration-by-iteration
                   agreement with textbook PCG.
* ``mas``        - invariants only (Galerkin transfer, per-domain inverse, SPD of the
                   applied operator); METIS partitions are not reproducible offline
                   (pymetis absent), so MAS domain choice is "parity unpinned".
"""
