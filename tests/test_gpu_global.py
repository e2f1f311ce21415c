"""Global updates (global_updates.jl:1-58) on the device against the oracle, through the C ABI:
same uniform stream, same decisions, same G and field afterwards; a rejected move restores
the chain exactly."""
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


def test_rejected_global_update_restores_state(amd):
    mod, DQMC = amd
    _say("global update with box 5 (rejected) L=4 M=20")
    from dqmc_amd import UniformStream
    p = config_params(mod, 4, 20)
    phi = np.random.default_rng(5).uniform(-1, 1, (3, p.N, p.slices))
    c = DQMC(p, uniforms=UniformStream(values=np.full(64, 0.999)))
    c.init(phi)
    G0, phi0, st0 = c.greens, c.hsfield, c.stack_element(1)
    S0 = c.stats().boson_action
    assert c.global_update(5.0) == 0
    st = c.stats()
    assert st.global_proposed == 1 and st.global_accepted == 0 and st.uniforms_used == 4
    assert np.array_equal(c.greens, G0) and np.array_equal(c.hsfield, phi0)
    assert all(np.array_equal(a, b) for a, b in zip(c.stack_element(1), st0))
    assert st.boson_action == S0 and (st.current_slice, st.direction) == (p.slices, -1)
    c.close()


def test_global_updates_match_oracle(amd):
    mod, DQMC = amd
    _say("update(mc, i) with global updates every 2nd ud-sweep, L=4 M=20, 4 ud-sweeps")
    from dqmc_amd import UniformStream
    p = config_params(mod, 4, 20)
    p.global_updates, p.global_rate, p.box_global = True, 2, 0.15
    seed = 41
    phi = np.random.default_rng(seed).uniform(-1, 1, (3, p.N, p.slices))
    c = DQMC(p, uniforms=UniformStream(seed + 100))
    c.set_proposal_log(8 * p.N * p.slices * 4 + 64)
    c.init(phi)
    o = O.DQMC(oracle_params(p), phi)
    o.init()
    ostream = O.UniformStream(O.host_uniforms(seed + 100, 16 * p.N * p.slices * 4 + 64))
    olog, oglob = [], []
    orig = o.global_update

    def spy(st, b):
        r = orig(st, b)
        oglob.append(r)
        return r

    o.global_update = spy
    for i in (1, 2, 3, 4):
        for _ in range(2 * p.slices):
            c.update(i)
            o.update(ostream, olog, i=i, global_rate=2, box_global=0.15)
    st = c.stats()
    assert st.global_proposed == len(oglob) == 2
    assert st.global_accepted == sum(oglob)
    assert _compare_logs(c.proposal_log(), olog) == 0
    assert c.stream.consumed == ostream.pos
    assert np.abs(c.hsfield - o.phi).max() < 1e-12
    assert np.abs(c.greens - o.greens).max() < 1e-10
    assert abs(st.boson_action - o.boson_action) < 1e-9 * abs(o.boson_action)
    c.close()
