"""CPU oracle for the C3L coupled-cluster hot path -- TEST INFRASTRUCTURE ONLY.

Nothing in the product package imports this directory.  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` use
it, always as the checker, never as the thing measured or shipped.

Modules
-------
``polynomial``  restatement of the reference's bit-mask polynomial residual and
                JVP (``src/c3l/cc/factorized.py:554-889``) plus the EOM sigma,
                sources and sink defined by ``src/c3l/cc/component_tensor.py``
                (``:2116-2302``, ``:2780-2858``), on sparse dict polynomials.
``fock``        sparse Fock-space brute force over all ``4n`` particle-hole
                modes (``exp(-T) K exp(T)`` by nilpotent series) and dense
                physical Lindblad exact diagonalisation (restating
                ``tests/reference/exact_generator.py:79-125`` and
                ``tests/cc/test_charge_s13.py:44-65,109-122``).
``port``        NumPy executor of the compiled contraction plans (the CPU port
                timed as the baseline beside the GPU).
``solvers``     FP64 CPU Newton--GMRES, shifted GMRES and Arnoldi resolvent
                restating ``cc/component_solver.py``, ``cc/component_response.py``
                and ``green/krylov.py`` on top of ``port``.

Parity pins are listed per function in each module header.
"""
