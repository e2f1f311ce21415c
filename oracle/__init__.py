"""CPU FP64 oracle for the DQMC sweep hot path of carstenbauer/dqmc.

TEST INFRASTRUCTURE ONLY. This package is the checker, never the product:
only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import it, and only to compute the expected answer or to time
the CPU baseline. The product path (``dqmc_amd``) never imports it and fails
loudly when its HIP library is missing.

The oracle restates the reference's algorithm operation for operation in
numpy, calling the same LAPACK routines the reference reaches through Julia's
``LinearAlgebra`` (``zgeqp3``, ``zungqr``, ``zgetrf``, ``zgetri``, ``zgetrs``),
here from scipy's bundled OpenBLAS. Every function cites the reference
``file:line`` it follows.

Parity pins (see ``tests/test_oracle.py``):
  * boson action, local action difference  -> reference KAT values
  * hopping / checkerboard / interaction operators -> O3.jld fixtures
  * slice matrices B, B^-1 (CBAssaad+BFIELD, CBFalse) -> O3.jld fixtures
  * calculate_greens / logdet -> O3.jld greens + Ul..Tr, direct calc_greens
  * calc_detratio / update_greens! / local sweep -> delta_i, M, det KAT,
    afterupdate_greens, afterlocal_greens, post-sweep boson action
  * decompose_udt / multiply_safely -> linalg.jld D_max / D_min KAT
  * stack propagation -> direct calc_greens at every slice (<1e-12, reference
    invariant tests_O3.jl:252-263), lambda=0 closed form
  * adjoint gradient -> central finite differences of S_B + logdet G
  * time-displaced G -> free-fermion closed form, reference identities
"""
