"""Pins the oracle's global update (global_updates.jl:1-58) by brute force.

The reference has no test of its own for global updates; the acceptance probability
is fixed by the mathematics: p = exp(-(S_B' - S_B)) det(1 + B'_M..B'_1) / det(1 + B_M..B_1)
with the shifted field, and the uniform stream is consumed 3 + (p <= 1) times.
"""
import numpy as np
import pytest

from oracle import dqmc_oracle as O


def _logdet_one_plus_chain(m):
    P = np.eye(m.p.n, dtype=np.complex128)
    for l in range(m.p.slices):
        P = m.multiply_B_left(l, P)
    return np.linalg.slogdet(np.eye(m.p.n) + P)[1]


@pytest.mark.parametrize("lam,seed", [(0.0, 1), (1.0, 2), (1.0, 3)])
def test_global_update_acceptance_brute_force(lam, seed):
    p = O.Params(4, 20, safe_mult=5, hoppings=(1.0, 0.5, 0.5, 1.0), lam=lam, chkr=True)
    rng = np.random.default_rng(seed)
    phi = rng.uniform(-1, 1, (3, p.N, p.slices))
    c = O.DQMC(p, phi)
    c.init()
    assert c.current_slice == p.slices and c.direction == -1
    u = rng.uniform(0, 1, 4)
    box = 0.3
    ld_old = _logdet_one_plus_chain(c.m)
    S_old = O.calc_boson_action(p, phi)
    shifted = phi + (-box + 2 * box * u[:3])[:, None, None]
    m_new = O.Model(p, shifted, c.m.ops)
    ld_new = _logdet_one_plus_chain(m_new)
    # log_det of the reference is log|det G| = -log|det(1 + B_M..B_1)|
    p_ref = np.exp(-(O.calc_boson_action(p, shifted) - S_old)) * np.exp(ld_new - ld_old)
    if lam == 0.0:
        assert abs(ld_old - ld_new) < 1e-10
    G_old, phi_old, stack_old = c.greens.copy(), c.phi.copy(), c.u_stack.copy()
    stream = O.UniformStream(u)
    acc = c.global_update(stream, box)
    p_acc, u_drawn = c.last_global
    assert abs(p_acc - p_ref) < 1e-9 * max(1.0, p_ref)
    assert stream.pos == (3 if p_acc > 1.0 else 4)
    assert acc == (1 if (p_acc > 1.0 or u[3] < p_acc) else 0)
    if acc:
        assert np.abs(c.phi - shifted).max() == 0.0
        assert np.abs(c.greens - O.calc_greens(m_new, p.slices)).max() < 1e-10
        assert abs(c.boson_action - O.calc_boson_action(p, shifted)) < 1e-12 * abs(c.boson_action)
    else:
        assert np.array_equal(c.phi, phi_old) and np.array_equal(c.greens, G_old)
        assert np.array_equal(c.u_stack, stack_old)
        assert c.boson_action == S_old


def test_update_with_global_rate_draw_order():
    # update(mc, i) (dqmc_framework.jl:500-517): one ud-sweep is 2M updates and lands on
    # (current_slice, direction) = (M, -1) once, at its last propagate; the global move
    # happens there when i % global_rate == 0, before that slice's local updates
    p = O.Params(4, 10, safe_mult=5, hoppings=(1.0, 0.5, 0.5, 1.0), lam=0.0, chkr=True)
    phi = np.random.default_rng(4).uniform(-1, 1, (3, p.N, p.slices))
    c = O.DQMC(p, phi)
    c.init()
    stream = O.UniformStream(O.host_uniforms(9, 4 * p.N * 2 * p.slices * 4 + 64))
    calls = []
    orig = c.global_update

    def spy(st, b):
        calls.append((c.current_slice, c.direction, st.pos))
        return orig(st, b)

    c.global_update = spy
    steps = 0
    for i in (1, 2, 3, 4):
        for _ in range(2 * p.slices):
            c.update(stream, None, i=i, global_rate=2, box_global=0.2)
            steps += 1
            if calls and len(calls) == (1 if i < 4 else 2) and steps % (2 * p.slices) == 0:
                assert calls[-1][:2] == (p.slices, -1)
    assert len(calls) == 2
