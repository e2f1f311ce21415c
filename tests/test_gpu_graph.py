"""Sweeps replayed from HIP graphs (dqmc_batch_use_graph / dqmc_use_graph): the same results
as eager launches, replay after replay, with the program's own eager work in between, and
the same decisions as the oracle."""
import sys

import numpy as np
import pytest

from tests.test_gpu_parity import _compare_logs, _run_pair, config_params

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


def test_graph_sweeps_match_eager_batch(amd):
    mod, DQMC = amd
    from dqmc_amd import Batch, UniformStream
    _say("batch of 3 chains (two shapes), graph replays vs eager, 5 sweeps")
    ps = [config_params(mod, 4, 20), config_params(mod, 6, 20), config_params(mod, 4, 20, lam=1.5)]
    rng = np.random.default_rng(71)
    confs = [rng.uniform(-1, 1, (3, p.N, p.slices)) for p in ps]
    bg = Batch(ps, streams=[UniformStream(700 + k) for k in range(3)])
    be = Batch(ps, streams=[UniformStream(700 + k) for k in range(3)])
    bg.use_graph(True)
    bg.init(confs)
    be.init(confs)
    for sweep in range(5):  # up, down, up (replay), down (replay), up (replay)
        bg.sweep(1)
        be.sweep(1)
        for k in range(len(ps)):
            sg, se = bg.stats(k), be.stats(k)
            assert np.abs(bg.greens(k) - be.greens(k)).max() == 0.0, (sweep, k)
            assert np.array_equal(bg.get_field(k), be.get_field(k)), (sweep, k)
            assert (sg.accepted, sg.uniforms_used, sg.current_slice, sg.direction) == \
                (se.accepted, se.uniforms_used, se.current_slice, se.direction), (sweep, k)
            assert sg.boson_action == se.boson_action and sg.drift_max < 1e-8
    bg.close()
    be.close()


def test_graph_chain_matches_oracle(amd):
    mod, DQMC = amd
    _say("single chain L=4 M=20 through sweep graphs vs oracle, 4 sweeps")
    p = config_params(mod, 4, 20)
    c, o, olog, _ = _run_pair(mod, DQMC, p, 73, 4, setup=lambda c: c.use_graph(True))
    assert _compare_logs(c.proposal_log(), olog) == 0
    assert np.abs(c.greens - o.greens).max() < 1e-10
    assert c.stats().drift_max < 1e-8
    c.close()
