"""Pins of the oracle's generic checkerboard (CBGeneric, hoppings_checkerboard_generic.jl,
slice_matrices.jl:233-393), CPU only.

The reference's CBGeneric fixtures (Bplus_CBGeneric, L=5) need its seeded configuration, which
only Julia's RNG reproduces, so the generic path is pinned by what fixes it mathematically:
1. the greedy bond groups partition the bonds, and no group touches a site twice;
2. B_l equals the dense product H_G^(1/2) ... H_2^(1/2) H_1 H_2^(1/2) ... H_G^(1/2) e^{dtau mu} eV_l
   built here directly from the group hopping matrices with scipy's expm (so the folding and
   the order of the factors are pinned), and each multiply_* variant is consistent with it
   (left/right, inverse, adjoint: tests_O3.jl:306-331);
3. the checkerboard error of prod_g H_g^(1/2) e^{dtau mu / 2} against exp(-dtau T / 2) is at most
   dtau^2 (check_chkr_quadratic, tests_O3.jl:72-91);
4. at lambda = 0 (free fermions) the stabilised stack gives G = (1 + B^M)^-1 exactly, B the
   field-independent checkerboard slice matrix;
5. for even L with the same groups forced, the Assaad and generic factorisations agree to
   O(dtau^2) (both approximate the same exp(-dtau T)).
A planted mistake in the order of the group factors (H_1 applied before the half steps)
fails check 2.
"""
import numpy as np
import pytest
from scipy.linalg import expm

from oracle import dqmc_oracle as O


def _params(L, M=10, lam=1.0, bfield=False, dt=0.1):
    return O.Params(L, M, delta_tau=dt, safe_mult=5, hoppings=(1.0, 0.5, 0.5, 1.0), lam=lam, chkr=True,
                    generic=True, bfield=bfield)


@pytest.mark.parametrize("L", [3, 5, 4])
def test_groups_partition_the_bonds(L):
    bonds = O.square_bonds(L)
    groups = O.build_checkerboard(L * L, bonds)
    ids = sorted(b for g in groups for b in g)
    assert ids == list(range(len(bonds)))
    for g in groups:
        sites = [s for b in g for s in bonds[b][:2]]
        assert len(sites) == len(set(sites))


def _dense_group_half(p, bonds, grp, sign, prefac):
    N = p.N
    blocks = []
    ph = O.peirls_phases(p) if p.bfield else None
    for (s_, f) in ((0, 0), (1, 1), (1, 0), (0, 1)):
        T = np.zeros((N, N), dtype=np.complex128)
        for b in grp:
            src, trg, v = bonds[b]
            t = p.t[1, f] if v == (0.0, 1.0) else p.t[0, f]
            pf = np.exp(1j * ph[s_][f][trg, src]) if p.bfield else 1.0
            pb = np.exp(1j * ph[s_][f][src, trg]) if p.bfield else 1.0
            T[trg, src] += -pf * t
            T[src, trg] += -pb * t
        blocks.append(expm(-sign * prefac * p.delta_tau * T))
    n = p.n
    out = np.zeros((n, n), dtype=np.complex128)
    for b, E in enumerate(blocks):
        out[b * N:(b + 1) * N, b * N:(b + 1) * N] = E
    return out


@pytest.mark.parametrize("L,bfield", [(3, False), (5, False), (5, True)])
def test_slice_matrix_equals_dense_product(L, bfield):
    p = _params(L, bfield=bfield)
    phi = np.random.default_rng(L).uniform(-1, 1, (3, p.N, p.slices))
    m = O.Model(p, phi)
    o = m.ops
    l = 3
    H = [_dense_group_half(p, o.bonds, g, 1.0, 0.5) for g in o.groups]
    H1 = _dense_group_half(p, o.bonds, o.groups[0], 1.0, 1.0)
    eV = O.interaction_matrix_exp(p, phi, l, 1.0).toarray()
    mu = o.chkr_mu.toarray()
    B = H1
    for Hg in H[1:]:
        B = Hg @ B @ Hg
    B = B @ mu @ eV
    eye = np.eye(p.n, dtype=np.complex128)
    Bl = m.multiply_B_left(l, eye)
    assert np.abs(Bl - B).max() < 1e-13
    assert np.abs(m.multiply_B_right(l, eye) - B).max() < 1e-13
    Binv = np.linalg.inv(B)
    assert np.abs(m.multiply_B_inv_left(l, eye) - Binv).max() < 1e-12
    assert np.abs(m.multiply_B_inv_right(l, eye) - Binv).max() < 1e-12
    A = np.random.default_rng(1).standard_normal((p.n, p.n)) + 0j
    assert np.abs(m.multiply_daggered_B_left(l, A) - B.conj().T @ A).max() < 1e-12
    # a planted ordering mistake (the full step before the half steps) is seen
    Bwrong = H1
    for Hg in H[1:]:
        Bwrong = Hg @ Hg @ Bwrong
    assert np.abs(Bl - Bwrong @ mu @ eV).max() > 1e-6


def test_checkerboard_error_quadratic():
    for dt in (0.1, 0.01, 0.001):
        p = _params(5, dt=dt)
        o = O.Operators(p)
        chk = np.eye(p.n)
        for Hh in o.chkr_hop_half:
            chk = Hh.toarray() @ chk
        chk = chk @ np.sqrt(o.chkr_mu.toarray())
        assert np.abs(o.hopping_matrix_exp - chk).max() <= dt ** 2


def test_free_fermion_stack_closed_form():
    p = _params(5, M=20, lam=0.0)
    phi = np.random.default_rng(2).uniform(-1, 1, (3, p.N, p.slices))
    d = O.DQMC(p, phi)
    d.init()
    B = d.m.multiply_B_left(0, np.eye(p.n, dtype=np.complex128))
    BM = np.linalg.matrix_power(B, p.slices)
    G = np.linalg.inv(np.eye(p.n) + BM)
    # init! leaves current_slice = M, direction -1: G(M) = [1 + B_{M-1}...B_1 B_M]^-1, the same for
    # a slice-independent B
    assert np.abs(d.greens - G).max() < 1e-10
    for _ in range(7):
        d.propagate()
    assert np.abs(d.greens - G).max() < 1e-10


def test_even_L_generic_close_to_assaad():
    pa = O.Params(4, 10, hoppings=(1.0, 0.5, 0.5, 1.0), chkr=True)
    pg = O.Params(4, 10, hoppings=(1.0, 0.5, 0.5, 1.0), chkr=True, generic=True)
    assert not pa.generic and pg.generic
    phi = np.random.default_rng(3).uniform(-1, 1, (3, pa.N, pa.slices))
    eye = np.eye(pa.n, dtype=np.complex128)
    Ba = O.Model(pa, phi).multiply_B_left(2, eye)
    Bg = O.Model(pg, phi).multiply_B_left(2, eye)
    assert np.abs(Ba - Bg).max() < 2 * pa.delta_tau ** 2
