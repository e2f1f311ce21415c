"""Merged (lockstep) sweeps of batches with several shapes against group-by-group sweeps and
the oracle, through the C ABI.

A merged sweep launches every kernel once for all chains of all shapes; each chain still
follows its own propagate state machine (stack.jl:391-499) through its plan rows. The same
uniform streams give the same decisions as sweeping the shape groups one by one: fields
identical, G within 1e-12 (the launches differ only in workgroup counts, i.e. summation
order), the same acceptance counts, stream positions and stack states.
"""
import sys

import numpy as np
import pytest

from oracle import dqmc_oracle as O
from tests.test_gpu_parity import config_params, oracle_params

pytestmark = pytest.mark.gpu


def _say(msg):
    print("[gpu-test] " + msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def amd():
    import dqmc_amd
    dqmc_amd.load()
    return dqmc_amd


def _batch(mod, ps, merged, seed0):
    from dqmc_amd import Batch, UniformStream
    b = Batch(ps, streams=[UniformStream(seed0 + k) for k in range(len(ps))])
    b.set_merged(merged)
    confs = [np.random.default_rng(seed0 + 50 + k).uniform(-1, 1, (3, p.N, p.slices)) for k, p in enumerate(ps)]
    b.init(confs)
    return b, confs


def test_merged_equals_groupwise(amd):
    mod = amd
    ps = [config_params(mod, 4, 20), config_params(mod, 6, 40, lam=1.5), config_params(mod, 4, 40, r=-1.0),
          config_params(mod, 6, 20, lam=0.7), config_params(mod, 8, 30, r=0.0), config_params(mod, 4, 20, lam=2.0)]
    _say("merged vs group-by-group: 6 chains, 5 shapes, 3 + 2 sweeps")
    bm, _ = _batch(mod, ps, True, 3100)
    bg, _ = _batch(mod, ps, False, 3100)
    for b in (bm, bg):
        b.sweep(3)
        b.sweep(2, sync=False)
        b.synchronize()
    for k, p in enumerate(ps):
        sm, sg = bm.stats(k), bg.stats(k)
        assert sm.device_error == 0
        assert (sm.current_slice, sm.direction) == (sg.current_slice, sg.direction), k
        assert sm.accepted == sg.accepted and sm.proposed == sg.proposed == 5 * p.N * p.slices, k
        assert bm.streams[k].consumed == bg.streams[k].consumed, k
        assert np.array_equal(bm.get_field(k), bg.get_field(k)), k
        assert np.abs(bm.greens(k) - bg.greens(k)).max() < 1e-12, k
        assert sm.drift_max < 1e-8 and abs(sm.drift_max - sg.drift_max) < 1e-12, k
        assert abs(sm.log_det - sg.log_det) < 1e-9 * max(1.0, abs(sg.log_det)), k
    bm.close()
    bg.close()


def test_merged_graph_replay_equals_eager(amd):
    """Merged sweeps replayed from HIP graphs (one graph per start state of the groups, each
    one sweep of every chain) against eager merged calls (each chain's sweeps back to back):
    every chain runs the same steps, but the set of chains sharing a launch differs, and with
    it the local-update windows, so G agrees to rounding (1e-12) while the fields, counters and
    stream positions are identical. The second call's uniform count changes the descriptors,
    which the cached graphs must pick up."""
    mod = amd
    ps = [config_params(mod, 4, 20), config_params(mod, 6, 40, lam=1.5), config_params(mod, 8, 20, r=0.0)]
    _say("merged graph replay vs eager merged: 3 shapes, 4 sweeps")
    bm, _ = _batch(mod, ps, True, 3500)
    bg, _ = _batch(mod, ps, True, 3500)
    bg.use_graph(True)
    for b in (bm, bg):
        b.sweep(1)
        b.sweep(3, sync=False)
        b.synchronize()
    for k, p in enumerate(ps):
        sm, sg = bm.stats(k), bg.stats(k)
        assert sg.device_error == 0
        assert (sm.current_slice, sm.direction) == (sg.current_slice, sg.direction), k
        assert sm.accepted == sg.accepted and sg.proposed == 4 * p.N * p.slices, k
        assert bm.streams[k].consumed == bg.streams[k].consumed, k
        assert np.array_equal(bm.get_field(k), bg.get_field(k)), k
        assert np.abs(bm.greens(k) - bg.greens(k)).max() < 1e-12, k
        assert sg.drift_max < 1e-8 and abs(sm.drift_max - sg.drift_max) < 1e-12, k
    bm.close()
    bg.close()


def test_merged_workgroups_by_size_equal_groupwise(amd):
    """16 chains of L=10 and L=12 (n = 400, 576): their columns of G do not fit in LDS at the
    merged launch's workgroup counts, so the merged local updates and QR give each chain
    workgroups by size (a block map over a 1-D cooperative grid) and each chain sizes its own
    window. Same decisions as sweeping the two shape groups one by one, G to rounding."""
    mod = amd
    ps = [config_params(mod, 10 if k % 2 == 0 else 12, 20, lam=0.8 + 0.05 * k) for k in range(16)]
    _say("merged, workgroups by size: 8 x L=10 and 8 x L=12, M=20, 2 sweeps")
    bm, _ = _batch(mod, ps, True, 3600)
    bg, _ = _batch(mod, ps, False, 3600)
    for b in (bm, bg):
        b.sweep(2)
    for k, p in enumerate(ps):
        sm, sg = bm.stats(k), bg.stats(k)
        assert sm.device_error == 0, k
        assert sm.accepted == sg.accepted and sm.proposed == sg.proposed == 2 * p.N * p.slices, k
        assert bm.streams[k].consumed == bg.streams[k].consumed, k
        assert np.array_equal(bm.get_field(k), bg.get_field(k)), k
        assert np.abs(bm.greens(k) - bg.greens(k)).max() < 1e-11, k
        assert sm.drift_max < 1e-8, k
    bm.close()
    bg.close()


def test_merged_chain_matches_oracle(amd):
    """A chain of a merged batch (L=6, beta=2, with L=4, beta=4 and L=8, beta=3 beside it)
    reproduces the oracle's field, G and stream consumption after 3 sweeps."""
    mod = amd
    ps = [config_params(mod, 4, 40), config_params(mod, 6, 20, lam=1.3), config_params(mod, 8, 30)]
    _say("merged batch chain vs oracle: L=6 M=20 next to L=4 M=40 and L=8 M=30, 3 sweeps")
    b, confs = _batch(mod, ps, True, 3200)
    b.sweep(3)
    k = 1
    p = ps[k]
    o = O.DQMC(oracle_params(p), confs[k])
    o.init()
    ostream = O.UniformStream(O.host_uniforms(3200 + k, 8 * p.N * p.slices * 3 + 64))
    for _ in range(3):
        o.sweep(ostream)
    assert np.abs(b.get_field(k) - o.phi).max() < 1e-12
    assert np.abs(b.greens(k) - o.greens).max() < 1e-10
    assert b.streams[k].consumed == ostream.pos
    st = b.stats(k)
    assert (st.current_slice, st.direction) == (o.current_slice, o.direction)
    b.close()


@pytest.mark.parametrize("L,M", [(6, 20), (12, 10)])
def test_wide_column_kernels_match_single_chains(amd, L, M):
    """From 8 chains per launch the QR solve and the Q formation run on the matrix cores with 16
    (n <= 400) or 8 (n <= 576) LDS-resident columns per workgroup, and the blocked QR's trailing
    update on the matrix cores where each trailing workgroup holds enough columns; a 9-chain
    batch must still equal single chains (the vector kernels)."""
    mod = amd
    from dqmc_amd import DQMC, Batch, UniformStream
    _say("matrix-core QR kernels: 9 chains L=%d M=%d vs single chains" % (L, M))
    ps = [config_params(mod, L, M, lam=0.8 + 0.1 * k) for k in range(9)]
    confs = [np.random.default_rng(3300 + k).uniform(-1, 1, (3, p.N, p.slices)) for k, p in enumerate(ps)]
    b = Batch(ps, streams=[UniformStream(3400 + k) for k in range(9)])
    b.init(confs)
    b.sweep(2)
    for k in (0, 4, 8):
        c = DQMC(ps[k], uniforms=UniformStream(3400 + k))
        c.init(confs[k])
        c.sweep(2)
        assert np.array_equal(b.get_field(k), c.hsfield), k
        assert np.abs(b.greens(k) - c.greens).max() < 1e-12, k
        c.close()
    b.close()
