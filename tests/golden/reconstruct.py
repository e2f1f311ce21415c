"""RNG-free reconstruction of the reference's seeded initial configuration.

The reference's local-update fixtures (test/tests_O3.jl:179-193) start from the
conf that init! draws from Julia's seeded global RNG (dqmc_framework.jl:152-155).
Slices 1-9 of that conf equal afterlocal_hsfield[:, :, 1:9] (the local sweep only
touches slice 10). Slice 10 is recovered from the Ur/Dr/Tr fixture, which holds
(B_10 ... B_1)^dagger at slice 1 going up (stack.jl:336):
    B_10 = (Ur Dr Tr)^dagger (B_9 ... B_1)^-1
    eV_10 = (chkr_hop_half[2] chkr_hop[1] chkr_hop_half[2] chkr_mu)^-1 B_10
and each site's phi follows from the eV block entries S and R
(interactions.jl:52-58): sinh(lambda dtau |phi|) = sqrt(|S|^2 + R^2).
"""
import math

import numpy as np

from oracle import dqmc_oracle as O

FIXTURE_PARAMS = dict(L=4, slices=10, delta_tau=0.1, safe_mult=10, hoppings=(1.0, 0.5, -0.5, -1.0),
                      mu1=-0.5, lam=0.5, r=2.0, c=3.0, u=1.0)


def fixture_params(chkr=True, bfield=True):
    """test/parameters/O3_generic_small_system.in.xml (and its no-bfield/no-chkr twins)."""
    return O.Params(chkr=chkr, bfield=bfield, **FIXTURE_PARAMS)


def reconstruct_seeded_conf(load):
    p = fixture_params()
    after = load("afterlocal_hsfield")
    m = O.Model(p, after.copy())
    Ur, Dr, Tr = load("Ur"), load("Dr"), load("Tr")
    X = ((Ur * Dr[None, :]) @ Tr).conj().T  # B_10 ... B_1
    for sl in range(9):
        X = m.multiply_B_inv_right(sl, X)
    o = m.ops
    # undo B_10 = H2h H1 H2h mu eV from the left, outermost factor first
    for op in (o.chkr_hop_half_inv[1], o.chkr_hop_inv[0], o.chkr_hop_half_inv[1], o.chkr_mu_inv):
        X = np.asarray(op @ X)
    N = p.N
    conf = after.copy()
    for i in range(N):
        S = X[i, N + i]
        R = X[i, 3 * N + i].real
        sinh_arg = math.sqrt(abs(S) ** 2 + R * R)
        nrm = math.asinh(sinh_arg) / (p.lam * p.delta_tau)
        sh = sinh_arg / nrm
        conf[:, i, 9] = [-S.real / sh, S.imag / sh, -R / sh]
    return conf
