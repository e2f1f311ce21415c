"""Pins the CPU oracle's time-displaced Green's functions and the effective -> true
Green's function transform.

The reference's own TDGF checks (tests_freefermions.jl:195-211,
tests_O3_measurements.jl:194-211) are identities and closed forms; they are restated
here at sizes where dense products are well conditioned, plus closed forms of the
reference's algorithm for a slice-independent B (free fermions, constant field).
"""
import numpy as np
import pytest

from oracle import dqmc_oracle as O
from oracle import tdgf as T


def _model(L, M, s, chkr, lam=1.0, seed=5, const=False):
    p = O.Params(L, M, safe_mult=s, hoppings=(1.0, 0.5, 0.5, 1.0), lam=lam, chkr=chkr)
    rng = np.random.default_rng(seed)
    if const:
        phi = np.repeat(rng.uniform(-1, 1, (3, p.N, 1)), M, axis=2)
    else:
        phi = rng.uniform(-1, 1, (3, p.N, M))
    return O.Model(p, phi)


def _prod(m, a, b):
    """Dense B(b) ... B(a), 1-based inclusive; identity when b < a."""
    P = np.eye(m.p.n, dtype=np.complex128)
    for k in range(a, b + 1):
        P = m.multiply_B_left(k - 1, P)
    return P


def _sym_slice(m, l):
    """The symmetric-Trotter slice matrix built from the operator pieces: hop_half[A]
    hop_half[B] mu eV hop_half[B] hop_half[A] (checkerboard) or eT eV eT (dense)."""
    o, p = m.ops, m.p
    eV = O.interaction_matrix_exp(p, m.phi, l, 1.0).toarray()
    if p.chkr:
        hA, hB = o.chkr_hop_half[0].toarray(), o.chkr_hop_half[1].toarray()
        return hA @ hB @ o.chkr_mu.toarray() @ eV @ hB @ hA
    eT = o.hopping_matrix_exp
    return eT @ eV @ eT


@pytest.mark.parametrize("chkr", [False, True])
def test_effective_greens_is_symmetric_trotter_greens(chkr):
    # effective_greens2greens! (fermion_measurements.jl:1125-1174) maps the G of the
    # asymmetric slice matrices to the G of the symmetric ones
    # L = 6: on the 4 x 4 torus the two plaquette groups commute, which would hide an
    # ordering mistake
    m = _model(6 if chkr else 4, 10, 5, chkr)
    M = m.p.slices
    for l in (1, 4, 10):
        G = O.calc_greens(m, l)
        P = np.eye(m.p.n, dtype=np.complex128)
        for k in list(range(l, M + 1)) + list(range(1, l)):
            P = _sym_slice(m, k - 1) @ P
        Gsym = np.linalg.inv(np.eye(m.p.n) + P)
        assert np.abs(T.effective_greens2greens(m, G) - Gsym).max() < 1e-10
    # U, T form (fermion_measurements.jl:1143-1159) agrees with the matrix form
    U, D, Tm = O.decompose_udt(G)
    U2, T2 = T.effective_greens2greens_ut(m, U, Tm)
    assert np.abs((U2 * D[None, :]) @ T2 - T.effective_greens2greens(m, G)).max() < 1e-12


def test_tdgf_free_fermions_closed_form():
    # lambda = 0, dense hopping: B = exp(-dtau T) for every slice, so
    # G(tau, 0) = B^(tau-1) (1 + B^M)^-1 and G(0, tau) = -(1 + B^M)^-1 B^(M - tau + 1)
    m = _model(4, 40, 10, False, lam=0.0)
    p = m.p
    e, V = np.linalg.eigh(m.ops.hopping_matrix)
    Gt0, G0t, st = T.measure_tdgfs(m)
    occ = 1.0 / (1.0 + np.exp(-p.beta * e))
    for tau in range(1, p.slices + 1):
        ref_t0 = (V * (np.exp(-(tau - 1) * p.delta_tau * e) * occ)) @ V.conj().T
        ref_0t = -(V * (np.exp(-(p.slices - tau + 1) * p.delta_tau * e) * occ)) @ V.conj().T
        assert np.abs(Gt0[tau - 1] - ref_t0).max() < 1e-10
        assert np.abs(G0t[tau - 1] - ref_0t).max() < 1e-10
    # tests_freefermions.jl:198-211
    assert np.abs(Gt0[0] - G0t[0] - np.eye(p.n)).max() < 1e-12
    for tau in range(2, p.slices + 1):
        assert np.abs(Gt0[tau - 1] + G0t[p.slices - tau + 1]).max() < 1e-10


def test_tdgf_constant_field_closed_form():
    # checkerboard, slice-independent field: the reference's recipe (checkpoint slices
    # from the UDT chains, B-propagation in between, fill_tdgf! :1513-1541) in closed form
    m = _model(6, 20, 5, True, const=True)
    p = m.p
    M, s = p.slices, p.safe_mult
    B = m.slice_matrix(0, 1.0)
    Bi = np.linalg.inv(B)
    mp = np.linalg.matrix_power
    G = np.linalg.inv(np.eye(p.n) + mp(B, M))
    E = lambda X: T.effective_greens2greens(m, X)  # noqa: E731
    Gt0, G0t, _ = T.measure_tdgfs(m)
    safe = list(range(1, M + 1, s))
    mhalf = M // 2 + 1
    for tau in range(1, M + 1):
        if tau in safe:
            ts, k = tau, 0
        elif tau >= mhalf:
            ts = max(x for x in safe if x < tau)
            k = tau - ts
        else:
            ts = min(x for x in safe if x > tau)
            k = -(ts - tau)
        Bk = mp(B, k) if k >= 0 else mp(Bi, -k)
        Bmk = mp(Bi, k) if k >= 0 else mp(B, -k)
        rt0 = Bk @ E(mp(B, ts - 1) @ G)
        r0t = -E(G @ mp(B, M - ts + 1)) @ Bmk
        assert np.abs(Gt0[tau - 1] - rt0).max() < 1e-10, tau
        assert np.abs(G0t[tau - 1] - r0t).max() < 1e-10, tau


@pytest.mark.parametrize("chkr", [False, True])
def test_tdgf_random_field(chkr):
    m = _model(6 if chkr else 4, 20, 5, chkr, seed=11)
    p = m.p
    M, s, n = p.slices, p.safe_mult, p.n
    Gt0, G0t, st = T.measure_tdgfs(m)
    eye = np.eye(n)
    # stacks: products of the slice matrices they stand for, and unitary U
    # (tests_O3_measurements.jl:200-211)
    for k in range(M // s):
        hi, lo = (k + 1) * s, k * s + 1
        want = {
            "BT0": _prod(m, 1, hi),
            "BT0Inv": np.linalg.inv(_prod(m, 1, hi)),
            "BBetaT": _prod(m, lo, M),
            "BBetaTInv": np.linalg.inv(_prod(m, lo, M)),
        }
        for name, W in want.items():
            U, D, Tm = (x[k] for x in st[name])
            assert np.abs(U @ U.conj().T - eye).max() < 1e-12
            assert np.abs((U * D[None, :]) @ Tm - W).max() < 1e-9 * max(1.0, np.abs(W).max())
    # checkpoint slices: G(tau, 1) = [B(tau-1, 1)^-1 + B(beta, tau)]^-1, effective -> true
    for tau in range(1, M + 1, s):
        direct = np.linalg.inv(np.linalg.inv(_prod(m, 1, tau - 1)) + _prod(m, tau, M))
        assert np.abs(Gt0[tau - 1] - T.effective_greens2greens(m, direct)).max() < 1e-10
        assert np.abs(Gt0[tau - 1] - T.calc_tdgf_direct(m, tau)).max() < 1e-10
        g0t = -np.linalg.inv(_prod(m, 1, tau - 1) + np.linalg.inv(_prod(m, tau, M)))
        assert np.abs(G0t[tau - 1] - T.effective_greens2greens(m, g0t)).max() < 1e-10
    # tests_O3_measurements.jl:196 and tests_freefermions.jl:198-201
    assert np.abs(Gt0[0] - T.effective_greens2greens(m, O.calc_greens(m, 1))).max() < 1e-10
    assert np.abs(Gt0[0] - G0t[0] - eye).max() < 1e-12
    # between checkpoints the reference propagates with B(tau) (upper half) and
    # B(tau+1)^-1 (lower half), fill_tdgf! :1522-1538
    assert np.abs(Gt0[M // 2 + 1] - m.multiply_B_left(M // 2 + 1, Gt0[M // 2])).max() < 1e-14
    assert np.abs(Gt0[M // 2 - 1] - m.multiply_B_inv_left(M // 2, Gt0[M // 2])).max() < 1e-14
