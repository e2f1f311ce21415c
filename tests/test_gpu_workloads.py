"""The benchmark workloads at their full sizes, through the C ABI.

* L=4, beta=5 dense (n=64, M=50): 10 sweeps with the oracle's accept/reject sequence.
* L=10, beta=20 checkerboard (n=400, M=200): 2 sweeps, identical decisions, G within 1e-10.
* The 64-chain mixed batch (L in {6,8,10,12}, beta in {5,10,20}, per-chain lambda and r):
  2 sweeps, drift below 1e-8 on every chain, and the L=6, beta=5 chains against the oracle.
* L=12, beta=10 (n=576, M=100, 43,200 components): dS/dphi against central differences of
  the device's own S, against the CPU trace formula on three slices, and HMC trajectories
  against a host leapfrog that uses the device gradient.
* L=16, beta=40 (n=1024, M=400): one sweep with drift below 1e-8 and G against the direct
  calc_greens at the end slice, the time-displaced Green's functions for all 400 slices
  (Gt0[1] - G0t[1] = 1, checkpoint slices against calc_tdgf_direct), one HMC trajectory.

Tolerances: G 1e-10 max-abs against the oracle; 1e-8 (the stated drift bound) for the
stabilised G against a fresh recompute on the device; dS/dphi 1e-6 relative against the
trace formula; finite differences within the error their step implies (see the test).
"""
import math
import sys
import time

import numpy as np
import pytest

from oracle import dqmc_oracle as O
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


def test_dense_L4_beta5_ten_sweeps(amd):
    """L=4, beta=5, no checkerboard (CBFalse): 10 sweeps, the oracle's decisions."""
    mod, DQMC = amd
    _say("L=4 M=50 dense, 10 sweeps vs oracle")
    p = config_params(mod, 4, 50, checkerboard=False)
    c, o, olog, _ = _run_pair(mod, DQMC, p, 1000, 10)
    assert _compare_logs(c.proposal_log(), olog) == 0
    assert np.abs(c.greens - o.greens).max() < 1e-10
    assert np.abs(c.hsfield - o.phi).max() < 1e-12
    assert c.stats().drift_max < 1e-8
    c.close()


def test_L10_beta20_two_sweeps(amd):
    """L=10, beta=20 (n=400, M=200), checkerboard: 2 sweeps with identical decisions and G
    within 1e-10 of the oracle after each (the oracle takes ~25 s per sweep on the host)."""
    mod, DQMC = amd
    _say("L=10 M=200, 2 sweeps vs oracle")
    p = config_params(mod, 10, 200)
    c, o, olog, _ = _run_pair(mod, DQMC, p, 1001, 2)
    assert _compare_logs(c.proposal_log(), olog) == 0
    assert np.abs(c.greens - o.greens).max() < 1e-10
    assert np.abs(c.hsfield - o.phi).max() < 1e-12
    st = c.stats()
    assert st.drift_max < 1e-8 and max(o.drifts) < 1e-8
    assert abs(st.boson_action - o.boson_action) < 1e-9 * abs(o.boson_action)
    c.close()


def mixed_batch_params(mod, nchains=64):
    """The mixed batch: chain c has L = [6,8,10,12][c mod 4], beta = [5,10,20][c//4 mod 3],
    r = [-1,0,1][c//12 mod 3], lambda = 0.5 + 1.5 u_c (u_c from seed 2000 + c)."""
    ps = []
    for c in range(nchains):
        L = (6, 8, 10, 12)[c % 4]
        beta = (5, 10, 20)[(c // 4) % 3]
        r = (-1.0, 0.0, 1.0)[(c // 12) % 3]
        lam = 0.5 + 1.5 * float(np.random.default_rng(2000 + c).uniform())
        p = config_params(mod, L, int(round(beta / 0.1)), lam=lam, r=r)
        p.seed = 2000 + c
        ps.append(p)
    return ps


def test_mixed_batch_64_chains(amd):
    """64 chains of 12 shapes advanced together for 2 sweeps; every chain's drift below 1e-8
    and G against a direct recompute; the six L=6, beta=5 chains (c = 0, 12, ..., 60) against
    the oracle: the same field after 2 sweeps (so the same decisions) and G within 1e-10."""
    mod, DQMC = amd
    from dqmc_amd import Batch, UniformStream
    ps = mixed_batch_params(mod)
    _say("mixed batch: 64 chains, 12 shapes, 2 sweeps")
    confs = [np.random.default_rng(p.seed).uniform(-1, 1, (3, p.N, p.slices)) for p in ps]
    b = Batch(ps, streams=[UniformStream(p.seed + 7) for p in ps])
    b.init(confs)
    b.sweep(2)
    for c, p in enumerate(ps):
        st = b.stats(c)
        assert st.device_error == 0, c
        assert st.drift_max < 1e-8, (c, st.drift_max)
        assert st.proposed == 2 * p.N * p.slices, c
    for c in (0, 12, 24, 36, 48, 60):
        p = ps[c]
        _say("mixed batch chain %d (L=%d M=%d r=%g lambda=%.4f) vs oracle" % (c, p.L, p.slices, p.r, p.lam))
        o = O.DQMC(oracle_params(p), confs[c])
        o.init()
        ostream = O.UniformStream(O.host_uniforms(p.seed + 7, 8 * p.N * p.slices * 2 + 64))
        for _ in range(2):
            o.sweep(ostream)
        assert np.abs(b.get_field(c) - o.phi).max() < 1e-12, c
        assert np.abs(b.greens(c) - o.greens).max() < 1e-10, c
        assert b.stats(c).uniforms_used == ostream.pos, c
    b.close()


def test_gradient_L12_beta10(amd):
    """L=12, beta=10 (n=576, M=100): dS/dphi for all 43,200 components.

    Central differences use the device's own S = S_B + log|det G| (dqmc_grad's S output, the
    stabilised logdet of calculate_logdet). With h = 1e-6 the difference carries a rounding
    error of about eps_S / h, eps_S ~ 1e-12 |S| from the logdet's backward error, plus
    h^2 |S'''| / 6, so the bound is 20 (1e-12 |S|) / h + 1e-7."""
    mod, DQMC = amd
    from dqmc_amd import UniformStream
    p = config_params(mod, 12, 100)
    rng = np.random.default_rng(3001)
    phi = rng.uniform(-1, 1, (3, p.N, p.slices))
    c = DQMC(p, uniforms=UniformStream(3002))
    c.init(phi)
    _say("gradient L=12 M=100")
    t0 = time.perf_counter()
    g, S = c.grad()
    _say("  one evaluation %.3f s, S = %.10g" % (time.perf_counter() - t0, S))
    assert g.shape == (3, p.N, p.slices) and np.isfinite(g).all()
    # S against the oracle's S_B and the device's direct logdet (calc_greens_and_logdet)
    _, ld1 = c.calc_greens_and_logdet(1)
    op = oracle_params(p)
    assert abs(S - (O.calc_boson_action(op, phi) + ld1)) < 1e-9 * abs(S)
    # the CPU FP64 trace formula on three slices (each costs a direct G on the host)
    sl = [0, 37, 99]
    g_o = O.action_grad_trace(O.Model(op, phi), slices=sl)
    for l in sl:
        err = np.abs(g[:, :, l] - g_o[:, :, l]).max() / max(1.0, np.abs(g_o[:, :, l]).max())
        assert err < 1e-6, (l, err)
    # central differences on 64 seeded components
    h = 1e-6
    tol = 20 * 1e-12 * abs(S) / h + 1e-7
    worst = 0.0
    for k in range(64):
        a, i, l = int(rng.integers(3)), int(rng.integers(p.N)), int(rng.integers(p.slices))
        vals = []
        for sgn in (1.0, -1.0):
            ph = phi.copy()
            ph[a, i, l] += sgn * h
            c.hsfield = ph
            vals.append(c.grad()[1])
        fd = (vals[0] - vals[1]) / (2 * h)
        worst = max(worst, abs(g[a, i, l] - fd))
        assert abs(g[a, i, l] - fd) < tol, (k, a, i, l, g[a, i, l], fd, tol)
    _say("  64 central differences: worst |g - fd| = %.3g (bound %.3g)" % (worst, tol))
    c.close()


def test_hmc_L12_beta10(amd):
    """HMC at L=12, beta=10: two trajectories of 10 leapfrog steps (eps 0.02). Each equals a
    host leapfrog driven by the device gradient (field within 1e-9, dH within 1e-6), and halving
    the step (20 steps of 0.01) cuts |dH| by about 4 (second-order integrator)."""
    mod, DQMC = amd
    from dqmc_amd import UniformStream
    p = config_params(mod, 12, 100)
    rng = np.random.default_rng(3003)
    phi = rng.uniform(-1, 1, (3, p.N, p.slices))
    c = DQMC(p, uniforms=UniformStream(3004))
    dhs = []
    for traj in range(2):
        mom = rng.standard_normal((3, p.N, p.slices))
        start = phi if traj == 0 else c.hsfield
        _say("HMC L=12 M=100 trajectory %d: device" % traj)
        c.init(start)
        acc, dh = c.hmc_trajectory(mom, 10, 0.02, 0.0)  # u = 0: always accepted
        assert acc and math.isfinite(dh)
        end = c.hsfield
        _say("HMC trajectory %d: host leapfrog with device gradients (dH %.4g)" % (traj, dh))
        c.init(start)
        g, S0 = c.grad()
        pm = mom.copy()
        H0 = S0 + 0.5 * np.sum(pm * pm)
        x = start.copy()
        pm -= 0.01 * g
        for s in range(10):
            x += 0.02 * pm
            c.hsfield = x
            g, S = c.grad()
            pm -= (0.02 if s < 9 else 0.01) * g
        H1 = S + 0.5 * np.sum(pm * pm)
        assert np.abs(end - x).max() < 1e-9, traj
        assert abs(dh - (H1 - H0)) < 1e-6 * max(1.0, abs(H0)), (traj, dh, H1 - H0)
        dhs.append(dh)
        c.init(end)
    # step halving on the first trajectory's start
    mom = np.random.default_rng(3003).standard_normal((3, p.N, p.slices))
    c.init(phi)
    _, dh1 = c.hmc_trajectory(mom, 10, 0.02, 0.0)
    c.init(phi)
    _, dh2 = c.hmc_trajectory(mom, 20, 0.01, 0.0)
    _say("step halving: dH %.4g -> %.4g" % (dh1, dh2))
    assert 0.12 < abs(dh2 / dh1) < 0.45, (dh1, dh2)
    c.close()


def test_L16_beta40_sweep_tdgf_hmc(amd):
    """L=16, beta=40 (n=1024, M=400), checkerboard, safe_mult 10."""
    mod, DQMC = amd
    from dqmc_amd import UniformStream
    p = config_params(mod, 16, 400)
    phi = np.random.default_rng(4001).uniform(-1, 1, (3, p.N, p.slices))
    c = DQMC(p, uniforms=UniformStream(4002))
    _say("L=16 M=400: init")
    c.init(phi)
    G0, _ = c.calc_greens_and_logdet(p.slices)
    assert np.abs(c.greens - G0).max() < 1e-8
    _say("L=16 M=400: one sweep")
    t0 = time.perf_counter()
    c.sweep(1)
    st = c.stats()
    _say("  sweep %.2f s, acceptance %.3f, drift_max %.3g" % (time.perf_counter() - t0,
                                                             st.accepted / max(st.proposed, 1), st.drift_max))
    assert st.device_error == 0
    assert st.proposed == p.N * p.slices
    assert 0.0 < st.drift_max < 1e-8  # recomputes happened (all_checks) and stayed in bounds
    assert (st.current_slice, st.direction) == (1, 1)
    Gd, ld = c.calc_greens_and_logdet(1)
    assert np.abs(c.greens - Gd).max() < 1e-8
    # time-displaced Green's functions of the swept field, every slice
    _say("L=16 M=400: measure_tdgfs (18.8 GB of device storage)")
    c.allocate_tdgfs()
    t0 = time.perf_counter()
    c.measure_tdgfs()
    _say("  measure_tdgfs %.2f s" % (time.perf_counter() - t0))
    a1, b1 = c.tdgf(1)
    assert np.abs(a1 - b1 - np.eye(p.n)).max() < 1e-10  # tests_freefermions.jl:198-201
    for tau in (1, 11, 201, 391):  # checkpoint slices: the direct form (fermion_measurements.jl:1553-1576)
        a, _ = c.tdgf(tau)
        Gdir = c.calc_tdgf_direct(tau)
        assert np.abs(a - Gdir).max() < 1e-8, tau
    for tau in (2, 100, 200, 201, 300, 400):
        a, b = c.tdgf(tau)
        assert np.isfinite(a).all() and np.isfinite(b).all() and np.abs(a).max() < 10 and np.abs(b).max() < 10, tau
    # the observables built on them (fermion_measurements.jl:500-923), all 400 slices
    t0 = time.perf_counter()
    c.measure_greens()
    zp, zm = c.measure_zfpc()
    cp, cm = c.measure_zfcdc()
    xx, yy, rho = c.measure_zfccc()
    _say("  zfpc + zfcdc + zfccc %.2f s, sfdensity %.6g" % (time.perf_counter() - t0, rho))
    for arr in (zp, zm, cp, cm, xx, yy):
        assert arr.shape == (16, 16) and np.isfinite(arr).all()
    assert math.isfinite(rho)
    c.deallocate_tdgfs()
    # one HMC trajectory (3 leapfrog steps of 0.02) from the swept field
    _say("L=16 M=400: one HMC trajectory, 3 leapfrog steps")
    mom = np.random.default_rng(4003).standard_normal((3, p.N, p.slices))
    acc, dh = c.hmc_trajectory(mom, 3, 0.02, 0.5)
    _say("  dH %.4g accepted %s" % (dh, acc))
    assert math.isfinite(dh) and abs(dh) < 50.0
    st = c.stats()
    assert st.device_error == 0 and (st.current_slice, st.direction) == (p.slices, -1)
    Gm, _ = c.calc_greens_and_logdet(p.slices)
    assert np.abs(c.greens - Gm).max() < 1e-8
    c.close()
