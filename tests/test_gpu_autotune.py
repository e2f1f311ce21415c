"""The autotuner (dqmc_autotune / dqmc_batch_autotune): it leaves the Monte Carlo state bit for
bit as it was, reports a valid choice, and the tuned chains keep the oracle's decisions."""
import sys

import numpy as np
import pytest

from oracle import dqmc_oracle as O
from tests.test_gpu_parity import _compare_logs, config_params, oracle_params

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


def _check_report(r, n):
    assert r["local_variant"] in (2, 3, 4) and 1 <= r["local_workgroups"] <= n // 2
    assert r["qr_panel"] in (8, 16) and 1 <= r["qr_workgroups"] <= n // 8
    assert r["local_ms"] > 0 and r["qr_ms"] > 0


def test_autotune_keeps_state_and_oracle_decisions(amd):
    mod, DQMC = amd
    from dqmc_amd import UniformStream
    _say("autotune L=6 M=20, then 3 sweeps vs oracle")
    p = config_params(mod, 6, 20)
    seed = 81
    phi = np.random.default_rng(seed).uniform(-1, 1, (3, p.N, p.slices))
    c = DQMC(p, uniforms=UniformStream(seed + 100))
    c.set_proposal_log(p.N * p.slices * 3 + 16)
    c.init(phi)
    G0, f0, s0 = c.greens, c.hsfield, c.stats()
    rep = c.autotune(reps=1)
    _check_report(rep, p.n)
    s1 = c.stats()
    assert np.array_equal(c.greens, G0) and np.array_equal(c.hsfield, f0)
    assert (s1.boson_action, s1.accepted, s1.proposed, s1.uniforms_used, s1.current_slice) == \
        (s0.boson_action, s0.accepted, s0.proposed, s0.uniforms_used, s0.current_slice)
    o = O.DQMC(oracle_params(p), phi)
    o.init()
    ostream = O.UniformStream(O.host_uniforms(seed + 100, 8 * p.N * p.slices * 3 + 64))
    olog = []
    for _ in range(3):
        c.sweep(1)
        o.sweep(ostream, olog)
    assert _compare_logs(c.proposal_log(), olog) == 0
    assert np.abs(c.greens - o.greens).max() < 1e-10
    c.close()


def test_batch_autotune_matches_untuned(amd):
    mod, DQMC = amd
    from dqmc_amd import Batch, UniformStream
    _say("batch of 3 chains (two shapes): autotuned vs default settings, 2 sweeps")
    ps = [config_params(mod, 4, 20), config_params(mod, 6, 20), config_params(mod, 4, 20, lam=1.5)]
    rng = np.random.default_rng(83)
    confs = [rng.uniform(-1, 1, (3, p.N, p.slices)) for p in ps]
    bt = Batch(ps, streams=[UniformStream(800 + k) for k in range(3)])
    bd = Batch(ps, streams=[UniformStream(800 + k) for k in range(3)])
    bt.init(confs)
    bd.init(confs)
    reps = bt.autotune(reps=1)
    assert len(reps) == 2
    _check_report(reps[0], ps[0].n)
    _check_report(reps[1], ps[1].n)
    bt.sweep(2)
    bd.sweep(2)
    for k in range(3):
        assert np.abs(bt.greens(k) - bd.greens(k)).max() < 1e-10
        assert bt.stats(k).accepted == bd.stats(k).accepted
    bt.close()
    bd.close()
