"""The generic (bond) checkerboard on the device, through the C ABI, against the oracle.

Odd L selects it (dqmc_framework.jl:109-115), as does forcing it at even L. Tolerances as in
test_gpu_parity: 1e-12 for single slice-matrix applications, 1e-10 for G after sweeps, the
oracle's accept/reject sequence exactly.
"""
import sys

import numpy as np
import pytest

from oracle import dqmc_oracle as O
from oracle import tdgf as OT
from tests.test_gpu_parity import _compare_logs, _run_pair, config_params, oracle_params

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


def _gparams(mod, L, M, force=False, bfield=False, lam=1.0, safe_mult=10):
    p = config_params(mod, L, M, lam=lam, safe_mult=safe_mult)
    p.generic_chkr = force
    p.bfield = bfield
    assert p.generic
    return p


@pytest.mark.parametrize("L,force,bfield", [(3, False, False), (5, False, False), (5, False, True), (4, True, False),
                                            (7, False, False)])
def test_slice_matrices_generic(amd, L, force, bfield):
    mod, DQMC = amd
    _say("generic checkerboard slice matrices L=%d force=%d bfield=%d" % (L, force, bfield))
    p = _gparams(mod, L, 10, force=force, bfield=bfield)
    phi = np.random.default_rng(L).uniform(-1, 1, (3, p.N, p.slices))
    c = DQMC(p)
    c.hsfield = phi
    m = O.Model(oracle_params(p), phi)
    X = np.random.default_rng(1).standard_normal((p.n, p.n)) + 1j * np.random.default_rng(2).standard_normal((p.n, p.n))
    for l in (0, 3, 9):
        for got, ref in ((c.multiply_B_left(l + 1, X), m.multiply_B_left(l, X)),
                         (c.multiply_B_right(l + 1, X), m.multiply_B_right(l, X)),
                         (c.multiply_B_inv_left(l + 1, X), m.multiply_B_inv_left(l, X)),
                         (c.multiply_B_inv_right(l + 1, X), m.multiply_B_inv_right(l, X)),
                         (c.multiply_daggered_B_left(l + 1, X), m.multiply_daggered_B_left(l, X))):
            assert np.abs(got - ref).max() < 1e-12 * max(1.0, np.abs(ref).max())
    assert np.abs(c.effective_greens2greens(X) - OT.effective_greens2greens(m, X)).max() < 1e-12 * np.abs(X).max()
    c.close()


@pytest.mark.parametrize("L,force", [(5, False), (4, True), (3, False)])
def test_seeded_chain_generic(amd, L, force):
    """3 sweeps of a seeded chain (beta 2) with the oracle's decisions and G within 1e-10."""
    mod, DQMC = amd
    _say("generic checkerboard chain L=%d M=20, 3 sweeps" % L)
    p = _gparams(mod, L, 20, force=force)
    c, o, olog, _ = _run_pair(mod, DQMC, p, 1200 + L, 3)
    assert _compare_logs(c.proposal_log(), olog) == 0
    assert np.abs(c.greens - o.greens).max() < 1e-10
    assert np.abs(c.hsfield - o.phi).max() < 1e-12
    assert c.stats().drift_max < 1e-8
    c.close()


def test_generic_larger_chain(amd):
    """L=9 (n=324), beta 5 (M=50): 1 sweep against the oracle."""
    mod, DQMC = amd
    _say("generic checkerboard chain L=9 M=50, 1 sweep")
    p = _gparams(mod, 9, 50)
    c, o, olog, _ = _run_pair(mod, DQMC, p, 1209, 1)
    assert _compare_logs(c.proposal_log(), olog) == 0
    assert np.abs(c.greens - o.greens).max() < 1e-10
    c.close()


def test_generic_free_fermion_and_tdgf(amd):
    """lambda = 0 at L=5: G = (1 + B^M)^-1 with the checkerboard B; time-displaced Green's
    functions against the oracle at lambda = 1."""
    mod, DQMC = amd
    _say("generic checkerboard L=5: free-fermion closed form and TDGF")
    p = _gparams(mod, 5, 20, lam=0.0, safe_mult=5)
    c = DQMC(p)
    c.init(np.random.default_rng(4).uniform(-1, 1, (3, p.N, p.slices)))
    B = c.multiply_B_left(1, np.eye(p.n, dtype=np.complex128))
    G = np.linalg.inv(np.eye(p.n) + np.linalg.matrix_power(B, p.slices))
    assert np.abs(c.greens - G).max() < 1e-10
    c.close()
    p = _gparams(mod, 5, 20, safe_mult=5)
    phi = np.random.default_rng(5).uniform(-1, 1, (3, p.N, p.slices))
    c = DQMC(p)
    c.hsfield = phi
    m = O.Model(oracle_params(p), phi)
    Gt0, G0t, _ = OT.measure_tdgfs(m)
    c.allocate_tdgfs()
    c.measure_tdgfs()
    for tau in (1, 2, 6, 11, 15, 20):
        a, b = c.tdgf(tau)
        assert np.abs(a - Gt0[tau - 1]).max() < 1e-10 and np.abs(b - G0t[tau - 1]).max() < 1e-10, tau
    c.close()
