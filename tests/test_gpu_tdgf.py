"""Parity of the HIP time-displaced Green's functions and of the effective -> true Green's
function transform with the CPU oracle (oracle/tdgf.py), through the C ABI.

Tolerances: 1e-10 max-abs for Green's functions (the stated bound for G), 1e-9 relative
for the checkpoint-chain products U D T (their entries span the chain's scales).
"""
import sys

import numpy as np
import pytest

from oracle import dqmc_oracle as O
from oracle import tdgf as OT
from tests.test_gpu_parity import config_params, oracle_params

pytestmark = pytest.mark.gpu


def _say(msg):
    print("[gpu-test] " + msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def amd():
    import dqmc_amd
    from dqmc_amd import DQMC
    dqmc_amd.load()
    return dqmc_amd, DQMC


def _pair(mod, DQMC, L, M, s, chkr, seed, lam=1.0):
    p = config_params(mod, L, M, checkerboard=chkr, safe_mult=s, lam=lam)
    phi = np.random.default_rng(seed).uniform(-1, 1, (3, p.N, p.slices))
    c = DQMC(p)
    c.hsfield = phi
    return c, O.Model(oracle_params(p), phi)


@pytest.mark.parametrize("chkr", [True, False])
def test_effective_greens2greens(amd, chkr):
    mod, DQMC = amd
    L = 6 if chkr else 4
    _say("effective_greens2greens L=%d chkr=%d" % (L, chkr))
    c, m = _pair(mod, DQMC, L, 10, 5, chkr, 3)
    rng = np.random.default_rng(4)
    G = rng.standard_normal((c.p.n, c.p.n)) + 1j * rng.standard_normal((c.p.n, c.p.n))
    assert np.abs(c.effective_greens2greens(G) - OT.effective_greens2greens(m, G)).max() < 1e-12
    c.close()


@pytest.mark.parametrize("chkr", [True, False])
def test_tdgf_matches_oracle(amd, chkr):
    mod, DQMC = amd
    L = 6 if chkr else 4
    _say("measure_tdgfs L=%d M=20 safe_mult=5 chkr=%d" % (L, chkr))
    c, m = _pair(mod, DQMC, L, 20, 5, chkr, 7)
    Gt0, G0t, st = OT.measure_tdgfs(m)
    c.allocate_tdgfs()
    c.measure_tdgfs()
    M, s = c.p.slices, c.p.safe_mult
    for tau in range(1, M + 1):
        a, b = c.tdgf(tau)
        assert np.abs(a - Gt0[tau - 1]).max() < 1e-10, ("Gt0", tau)
        assert np.abs(b - G0t[tau - 1]).max() < 1e-10, ("G0t", tau)
    eye = np.eye(c.p.n)
    for name in c.TDGF_STACKS:
        for k in range(M // s):
            U, D, T = c.tdgf_stack(name, k)
            Uo, Do, To = (x[k] for x in st[name])
            W = (Uo * Do[None, :]) @ To
            assert np.abs(U @ U.conj().T - eye).max() < 1e-12, (name, k)
            assert np.abs((U * D[None, :]) @ T - W).max() < 1e-9 * np.abs(W).max(), (name, k)
            assert np.abs(D - Do).max() < 1e-9 * Do.max(), (name, k)
    # a second measurement after changing the field reuses the storage
    phi2 = np.random.default_rng(8).uniform(-1, 1, (3, c.p.N, M))
    c.hsfield = phi2
    m2 = O.Model(m.p, phi2, m.ops)
    Gt0b, G0tb, _ = OT.measure_tdgfs(m2)
    c.measure_tdgfs()
    for tau in (1, 2, M // 2, M // 2 + 1, M):
        a, b = c.tdgf(tau)
        assert np.abs(a - Gt0b[tau - 1]).max() < 1e-10 and np.abs(b - G0tb[tau - 1]).max() < 1e-10, tau
    c.deallocate_tdgfs()
    c.close()


@pytest.mark.parametrize("chkr", [True, False])
def test_calc_tdgf_direct(amd, chkr):
    mod, DQMC = amd
    L = 6 if chkr else 4
    _say("calc_tdgf_direct L=%d M=20 chkr=%d" % (L, chkr))
    c, m = _pair(mod, DQMC, L, 20, 5, chkr, 9)
    for sl in (1, 4, 11, 20):
        assert np.abs(c.calc_tdgf_direct(sl) - OT.calc_tdgf_direct(m, sl)).max() < 1e-10, sl
    c.close()


def test_tdgf_free_fermions_large(amd):
    """lambda = 0 without checkerboard at L=8, beta=10 (n=256, M=100): the closed form
    G(tau, 0) = B^(tau-1) (1 + B^M)^-1, G(0, tau) = -(1 + B^M)^-1 B^(M-tau+1), B = exp(-dtau T),
    and G(tau, 0) = -G(0, beta - tau) (tests_freefermions.jl:195-211)."""
    mod, DQMC = amd
    _say("free-fermion TDGF L=8 M=100 dense")
    c, m = _pair(mod, DQMC, 8, 100, 10, False, 1, lam=0.0)
    c.allocate_tdgfs()
    c.measure_tdgfs()
    e, V = np.linalg.eigh(m.ops.hopping_matrix)
    p = m.p
    occ = 1.0 / (1.0 + np.exp(-p.beta * e))
    g0 = {}
    for tau in range(1, p.slices + 1):
        a, b = c.tdgf(tau)
        ref_t0 = (V * (np.exp(-(tau - 1) * p.delta_tau * e) * occ)) @ V.conj().T
        ref_0t = -(V * (np.exp(-(p.slices - tau + 1) * p.delta_tau * e) * occ)) @ V.conj().T
        assert np.abs(a - ref_t0).max() < 1e-10, tau
        assert np.abs(b - ref_0t).max() < 1e-10, tau
        g0[tau] = (a, b)
    assert np.abs(g0[1][0] - g0[1][1] - np.eye(p.n)).max() < 1e-12
    for tau in range(2, p.slices + 1):
        assert np.abs(g0[tau][0] + g0[p.slices - tau + 2][1]).max() < 1e-10
    c.close()
