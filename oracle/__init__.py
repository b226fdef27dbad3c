"""CPU oracle for the split-operator propagators: TEST INFRASTRUCTURE ONLY.

This package restates, in NumPy complex128, the operator sequences of the reference
propagators (wigner_moyal_cuda_1d.py, wigner_bloch_cuda_1d.py, rho_vneumann_cuda_1d.py,
rho_bloch_cuda_1d.py, schrodinger_wigner_cuda_1d.py, with cufft.py's axis/sign
conventions).  Each function cites the reference file:line it follows.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py``
may import it, and only as the checker.  The product package (``accuratewigner_amd``)
never imports it and has no CPU fallback.

Pinning (which test fixes which function) is listed in ``oracle/awig_oracle.py``'s module
docstring.
"""
