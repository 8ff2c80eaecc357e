"""CPU FP64 oracle for the FFT-VIE solve path.

TEST INFRASTRUCTURE ONLY. This package is a from-scratch NumPy restatement of
VoxPhoton's MATLAB algorithms for the N-operator matvec, the Chan circulant
preconditioners, restarted GMRES and the Green-tensor face integrals.  Only
``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import it, and only as the checker (or the timed CPU
baseline).  The product path (``voxphoton_amd``) never imports it.

Each function cites the reference file:line it restates (paths are relative to
the reference repository root; ``pc/`` = ``piecewise_constant/``).

Pinning (what checks each function against something other than itself):
  * ``vie.mv_AN_const``         dense brute-force Toeplitz product on tiny grids
                                (closed-form identity of circulant embedding),
                                adjoint identity, Mc=0 => Gram*Mr*I, linearity.
  * ``vie.circulant_nop_const`` via the dense product above (parity signs).
  * ``precond.*``               dense F^-1 blkdiag(B) F products; exactly
                                circulant operators give the identity; the
                                direct 3Nx3N block construction against the
                                literal 3MNx3MN restatement; Chan optimality
                                against least squares.
  * ``krylov.pgmres``           numpy.linalg.solve on small dense systems,
                                Arnoldi relation, n-step termination.
  * ``greens``                  static closed form of the coincident-square
                                1/R integral, order-doubling convergence,
                                brute-force tensor Gauss on far cells.
"""
