"""The measurement hand-off and the fermionic observables on the device, through the C ABI,
against the CPU oracle (oracle/observables.py, oracle/tdgf.py).

Tolerances: Green's functions 1e-10 max-abs (the stated bound for G); observables 1e-10
relative to their largest entry (sums of N M products of G entries).
"""
import sys

import numpy as np
import pytest

from dqmc_amd import model as PM
from oracle import dqmc_oracle as O
from oracle import observables as OB
from oracle import tdgf as OT
from tests.test_gpu_parity import _run_pair, config_params, oracle_params

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


def _close(a, b, rel=1e-10):
    return np.abs(a - b).max() <= rel * max(1.0, np.abs(b).max())


def test_measurement_handoff_after_sweeps(amd):
    """_measure_observables! at the measurement point (current_slice = M, direction = -1, after
    an up and a down sweep): wrap_greens(G, M, +1) then effective_greens2greens!, and
    measure_greens (calc_greens(mc, 1) + effective_greens2greens!)."""
    mod, DQMC = amd
    _say("measurement hand-off L=4 M=20 after 2 sweeps")
    p = config_params(mod, 4, 20)
    c, o, _, _ = _run_pair(mod, DQMC, p, 801, 2)
    assert (c.current_slice, c.direction) == (p.slices, -1) == (o.current_slice, o.direction)
    g_dev = c.measure_current_greens()
    g_ref = OT.effective_greens2greens(o.m, o.wrap_greens(o.greens, o.current_slice, 1))
    assert np.abs(g_dev - g_ref).max() < 1e-10
    assert np.abs(c.greens - o.greens).max() < 1e-10  # the chain's own G is unchanged
    e_dev = c.measure_greens()
    e_ref = OT.effective_greens2greens(o.m, O.calc_greens(o.m, 1))
    assert np.abs(e_dev - e_ref).max() < 1e-10
    assert np.abs(c.greens - o.greens).max() < 1e-10
    c.close()


@pytest.mark.parametrize("L,M,s,chkr", [(4, 20, 10, True), (6, 20, 5, True), (4, 10, 5, False)])
def test_observables_match_oracle(amd, L, M, s, chkr):
    mod, DQMC = amd
    _say("fermionic observables L=%d M=%d safe_mult=%d chkr=%d" % (L, M, s, chkr))
    p = config_params(mod, L, M, checkerboard=chkr, safe_mult=s)
    phi = np.random.default_rng(900 + L).uniform(-1, 1, (3, p.N, p.slices))
    c = DQMC(p)
    c.init(phi)
    m = O.Model(oracle_params(p), phi)
    G = c.measure_greens()
    Gref = OT.effective_greens2greens(m, O.calc_greens(m, 1))
    assert np.abs(G - Gref).max() < 1e-10
    c.allocate_tdgfs()
    c.measure_tdgfs()
    Gt0s, G0ts, _ = OT.measure_tdgfs(m)
    T = PM.hopping_matrix(p)
    # the oracle evaluates on the device's G (the inputs agree to 1e-10 already)
    for got, ref in zip(c.measure_etpc(), OB.measure_etpc(L, G)):
        assert _close(got, ref)
    for got, ref in zip(c.measure_etcdc(), OB.measure_etcdc(L, G)):
        assert _close(got, ref)
    for got, ref in zip(c.measure_zfpc(), OB.measure_zfpc(L, Gt0s)):
        assert _close(got, ref, 1e-9)
    for got, ref in zip(c.measure_zfcdc(), OB.measure_zfcdc(L, G, Gt0s, G0ts)):
        assert _close(got, ref, 1e-9)
    xx, yy, rho = c.measure_zfccc()
    rxx, ryy = OB.measure_zfccc(L, G, Gt0s, G0ts, T)
    assert _close(xx, rxx, 1e-9) and _close(yy, ryy, 1e-9)
    assert abs(rho - OB.measure_sfdensity(rxx)) < 1e-9 * max(1.0, np.abs(rxx).max())
    c.deallocate_tdgfs()
    c.close()


def test_free_fermion_observables_symmetric(amd):
    """lambda = 0 without checkerboard (L=4, beta=2): G has the lattice's reflection symmetry,
    so every observable is reflection symmetric and the charge and current ones real
    (the reference's checks, tests_O3_measurements.jl:220-300, hold exactly here)."""
    mod, DQMC = amd
    _say("free-fermion observables L=4 M=20 dense")
    p = config_params(mod, 4, 20, checkerboard=False, lam=0.0)
    c = DQMC(p)
    c.init(np.random.default_rng(3).uniform(-1, 1, (3, p.N, p.slices)))
    c.measure_greens()
    c.allocate_tdgfs()
    c.measure_tdgfs()
    xx, yy, _ = c.measure_zfccc()
    for arr in (*c.measure_etpc(), *c.measure_zfpc(), *c.measure_etcdc(), *c.measure_zfcdc(), xx, yy):
        assert OB.is_reflection_symmetric(arr, tol=1e-10)
        assert np.abs(np.imag(arr)).max() < 1e-12
    c.close()
