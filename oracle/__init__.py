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
  * ``precond.*``               the 1- and 2-level block producers equal the
                                orthogonal (Frobenius) projection of the dense
                                operator onto x- / xy-block-circulant matrices;
                                exactly circulant operators are inverted;
                                dense F^-1 blkdiag(B) F products for both
                                applies; chan_x against least squares; the
                                composite against its two parts.
  * ``krylov.pgmres``           numpy.linalg.solve on small dense systems,
                                Arnoldi relation, n-step termination.
  * ``greens``                  static closed form of the coincident-square
                                1/R integral, order-doubling convergence,
                                brute-force tensor Gauss on far cells, and the
                                reference's own DEMCEM singular values and
                                tensor cubature (``ref_greens``).
  * ``ref_greens``              the reference's C++ cubature cores compiled
                                from the reference sources into ``_ref/`` by
                                ``build_ref.sh`` (never on the product path).
"""
