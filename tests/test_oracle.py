"""Pins the CPU oracle against the reference's own fixtures and known-answer tests.

Each check names the reference test it mirrors. Tolerances follow the reference's
`isapprox` (rtol = sqrt(eps)) where it uses one; matrix fixtures are compared at
max-abs 1e-12 (they reproduce at ~1e-15).
"""
import math

import numpy as np
import pytest

from oracle import dqmc_oracle as O
from tests.conftest import load_fixture as ld
from tests.golden.reconstruct import fixture_params, reconstruct_seeded_conf

RTOL = math.sqrt(np.finfo(float).eps)


def close(a, b, rtol=RTOL):
    return abs(a - b) <= rtol * max(abs(a), abs(b))


def test_boson_action_kat():
    # tests_O3.jl:51-57
    p = fixture_params()
    conf = ld("randconf")
    m = O.Model(p, conf)
    assert close(m.boson_action(), 75.57712964980982)
    assert close(m.boson_action_diff(2, 9, [0.370422, 0.797014, 0.956094]), 0.015453643701304112)
    p.edrun = True
    assert close(m.boson_action(), 16.348917129437076)
    assert close(m.boson_action_diff(2, 9, [0.370422, 0.797014, 0.956094]), 0.014491910524600016)


def test_boson_action_diff_is_difference():
    # calc_boson_action_diff must equal S_B(new) - S_B(old) (action.jl:1-101)
    p = fixture_params()
    rng = np.random.default_rng(5)
    conf = rng.uniform(-1, 1, (3, p.N, p.slices))
    m = O.Model(p, conf)
    for site, sl in [(0, 0), (5, 9), (15, 4)]:
        new = conf[:, site, sl] + rng.uniform(-0.5, 0.5, 3)
        c2 = conf.copy()
        c2[:, site, sl] = new
        ref = O.calc_boson_action(p, c2) - O.calc_boson_action(p, conf)
        assert abs(m.boson_action_diff(site, sl, new) - ref) < 1e-12


def test_lattice_tables():
    # tests_O3.jl:267-272
    nb = O.neighbors(4) + 1
    assert nb.tolist() == [[2, 3, 4, 1, 6, 7, 8, 5, 10, 11, 12, 9, 14, 15, 16, 13],
                           [5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 1, 2, 3, 4],
                           [4, 1, 2, 3, 8, 5, 6, 7, 12, 9, 10, 11, 16, 13, 14, 15],
                           [13, 14, 15, 16, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12]]
    assert (O.time_neighbors(10) + 1).tolist() == [[2, 3, 4, 5, 6, 7, 8, 9, 10, 1], [10, 1, 2, 3, 4, 5, 6, 7, 8, 9]]
    A, B = O.four_site_corners(4)
    assert (A + 1).tolist() == [1, 3, 9, 11] and (B + 1).tolist() == [6, 8, 14, 16]  # tests_O3.jl:96


def test_hopping_matrices():
    # tests_O3.jl:62-66 (BFIELD=TRUE model), tests_O3_peirls.jl
    p = fixture_params(bfield=True)
    ops = O.Operators(p)
    assert np.abs(ops.hopping_matrix_exp - ld("hopping_matrix_exp")).max() < 1e-12
    assert np.abs(ops.hopping_matrix_exp_inv - ld("hopping_matrix_exp_inv")).max() < 1e-12
    ph = O.peirls_phases(p)
    for s in range(2):
        for f in range(2):
            ref = ld("peirls%d%d" % (s + 1, f + 1))
            assert np.array_equal(np.isnan(ph[s][f]), np.isnan(ref))
            assert np.nanmax(np.abs(ph[s][f] - ref)) < 1e-14


@pytest.mark.parametrize("bfield", [False, True])
def test_checkerboard_matrices(bfield):
    # tests_O3.jl:94-136
    p = fixture_params(bfield=bfield)
    ops = O.Operators(p)
    pre = "" if bfield else "nob_"
    pairs = [("chkr_hop1", ops.chkr_hop[0]), ("chkr_hop2", ops.chkr_hop[1]),
             ("chkr_hop_inv1", ops.chkr_hop_inv[0]), ("chkr_hop_inv2", ops.chkr_hop_inv[1]),
             ("chkr_hop_half1", ops.chkr_hop_half[0]), ("chkr_hop_half2", ops.chkr_hop_half[1]),
             ("chkr_hop_half_inv1", ops.chkr_hop_half_inv[0]), ("chkr_hop_half_inv2", ops.chkr_hop_half_inv[1]),
             ("chkr_mu", ops.chkr_mu), ("chkr_mu_inv", ops.chkr_mu_inv)]
    for name, mat in pairs:
        assert np.abs(mat.toarray() - ld(pre + name)).max() < 1e-12, name


@pytest.mark.parametrize("bfield", [False, True])
def test_checkerboard_trotter_error(bfield):
    # check_chkr_quadratic, tests_O3.jl:75-91: |e^{-dtau T/2} - H_A^half H_B^half mu^half| <= dtau^2
    for dt in (0.1, 0.01, 0.001):
        p = fixture_params(bfield=bfield)
        p.delta_tau = dt
        ops = O.Operators(p)
        chk = (ops.chkr_hop_half[0] @ ops.chkr_hop_half[1] @ ops.chkr_mu_half).toarray()
        assert np.abs(ops.hopping_matrix_exp - chk).max() <= dt ** 2


def test_interaction_exponentials():
    # tests_O3.jl:155-175; the BFIELD eV fixture was dumped with the seeded conf (slice 3)
    p = fixture_params(bfield=False)
    conf = ld("randconf")
    assert np.abs(O.interaction_matrix_exp(p, conf, 2, 1.0).toarray() - ld("nob_eVplus")).max() < 1e-14
    assert np.abs(O.interaction_matrix_exp(p, conf, 2, -1.0).toarray() - ld("nob_eVminus")).max() < 1e-14
    after = ld("afterlocal_hsfield")
    assert np.abs(O.interaction_matrix_exp(p, after, 2, 1.0).toarray() - ld("eVplus")).max() < 1e-14
    assert np.abs(O.interaction_matrix_exp(p, after, 2, -1.0).toarray() - ld("eVminus")).max() < 1e-14
    op = [0.130018, 0.792039, 0.683411]
    assert np.abs(O.interaction_matrix_exp_op(p, op, 1.0) - ld("nob_eVexpop")).max() < 1e-15
    assert np.abs(O.interaction_matrix_exp_op(p, op, 1.0) - ld("eVexpop")).max() < 1e-15
    # eV(+1) eV(-1) = I (exact inverse pair)
    e = O.interaction_matrix_exp(p, conf, 4, 1.0) @ O.interaction_matrix_exp(p, conf, 4, -1.0)
    assert np.abs(e.toarray() - np.eye(p.n)).max() < 1e-14


@pytest.mark.parametrize("chkr,bfield,suffix", [(True, True, ""), (False, False, "_CBFalse")])
def test_slice_matrices(chkr, bfield, suffix):
    # tests_O3.jl:283-331; slice 3 of the seeded conf equals afterlocal_hsfield slice 3
    p = fixture_params(chkr=chkr, bfield=bfield)
    m = O.Model(p, ld("afterlocal_hsfield"))
    B = m.slice_matrix(2, 1.0)
    Binv = m.slice_matrix(2, -1.0)
    assert np.abs(B - ld("Bplus" + suffix)).max() < 1e-12
    assert np.abs(Binv - ld("Bminus" + suffix)).max() < 1e-12
    rng = np.random.default_rng(0)
    A = rng.random((p.n, p.n)) + 1j * rng.random((p.n, p.n))
    assert np.abs(m.multiply_daggered_B_left(2, A) - B.conj().T @ A).max() < 1e-12
    assert np.abs(m.multiply_B_right(2, A) - A @ B).max() < 1e-12
    assert np.abs(m.multiply_B_inv_right(2, A) - A @ Binv).max() < 1e-12
    assert np.abs(m.multiply_B_inv_left(2, m.multiply_B_left(2, A)) - A).max() < 1e-12


def test_decompose_udt_invariants():
    # tests_linalg.jl:37-60 (D values need Julia's RNG input; invariants pinned instead)
    rng = np.random.default_rng(1234)
    x = rng.random((10, 10)).astype(np.complex128)
    U, D, T = O.decompose_udt(x)
    assert np.abs(U @ U.conj().T - np.eye(10)).max() < 1e-14
    assert sorted(int(np.sum(np.abs(T[:, j]) == 0)) for j in range(10)) == list(range(10))
    assert np.abs((U * D[None, :]) @ T - x).max() < 1e-14
    assert np.all(np.diff(D) <= 1e-15)  # column pivoting orders |R_ii|
    s = np.linalg.svd(x, compute_uv=False)
    assert close(D[0] * 1.0, D[0]) and D[0] <= s[0] * (1 + 1e-14) and D[-1] >= s[-1] * (1 - 1e-14)


def test_multiply_safely_kat():
    # tests_linalg.jl:105-113
    u = ld("B_QR_U").astype(np.complex128)
    d = ld("B_QR_D")
    t = ld("B_QR_T").astype(np.complex128)
    _, D, _ = O.multiply_safely(u, d, t, u, d, t)
    assert close(D.max(), 5.8846316709257896e16)
    assert close(D.min(), 1.9123571535539083e-24)


def test_calculate_greens_kat():
    # tests_O3.jl:215-230
    p = fixture_params()
    conf = reconstruct_seeded_conf(ld)
    c = O.DQMC(p, conf)
    c.Ul, c.Dl, c.Tl = ld("Ul"), ld("Dl"), ld("Tl")
    c.Ur, c.Dr, c.Tr = ld("Ur"), ld("Dr"), ld("Tr")
    c.calculate_greens()
    ldet = c.calculate_logdet()
    g, ldfresh = O.calc_greens_and_logdet(c.m, 1)
    assert np.abs(c.greens - ld("greens")).max() < 1e-12
    assert np.abs(c.greens - g).max() < 1e-12
    assert close(ldet, ldfresh, 1e-12)


def test_seeded_conf_reconstruction():
    # the 5 rejected sites of slice 10 keep their value through the sweep, so the
    # reconstruction must reproduce them; 11 changed sites = acceptance 0.6875 (tests_O3.jl:190)
    conf = reconstruct_seeded_conf(ld)
    after = ld("afterlocal_hsfield")
    diff = np.abs(conf[:, :, 9] - after[:, :, 9]).max(axis=0)
    assert np.sum(diff > 1e-9) == 11
    assert diff[diff <= 1e-9].max() < 1e-11
    assert np.array_equal(conf[:, :, :9], after[:, :, :9])


def test_local_update_kats():
    # tests_O3.jl:179-193
    p = fixture_params()
    conf = reconstruct_seeded_conf(ld)
    after = ld("afterlocal_hsfield")
    c = O.DQMC(p, conf.copy())
    c.init()
    assert (c.current_slice, c.direction) == (10, -1)
    det = c.calc_detratio(6, np.array([0.488033, 0.0196912, 0.438309]))
    assert abs(det - (1.000380293015979 + 1.0842021724855044e-18j)) < 1e-12
    assert np.abs(c.delta_i - ld("delta_i")).max() < 1e-12
    assert np.abs(c.M - ld("M")).max() < 1e-12
    c.update_greens(6)
    assert np.abs(c.greens - ld("afterupdate_greens")).max() < 1e-12
    # the sweep that followed: accepted sites are those whose phi changed
    for i in range(p.N):
        new = after[:, i, 9]
        if np.abs(new - conf[:, i, 9]).max() > 1e-9:
            exp_dS = math.exp(-c.m.boson_action_diff(i, 9, new))
            r = c.calc_detratio(i, new)
            assert r.real > 0
            c.phi[:, i, 9] = new
            c.boson_action += -math.log(exp_dS)
            c.update_greens(i)
    assert np.abs(c.greens - ld("afterlocal_greens")).max() < 1e-11
    assert close(c.boson_action, 75.18407422927604)


def test_propagation_matches_direct_greens():
    # tests_O3.jl:240-263: every slice of a full up-down cycle within 1e-12 of direct G
    p = fixture_params()
    c = O.DQMC(p, reconstruct_seeded_conf(ld))
    c.init()
    c.propagate()
    steps = 0
    while not (c.current_slice == p.slices and c.direction == -1):
        g = O.calc_greens(c.m, c.current_slice)
        assert np.abs(g - c.greens).max() < 1e-12
        c.propagate()
        steps += 1
    assert steps == 2 * p.slices - 1
    c2 = O.DQMC(p, reconstruct_seeded_conf(ld))
    c2.init()
    while c2.current_slice != 1:
        c2.propagate()
    c2.propagate()
    assert np.abs(c2.greens - ld("greens")).max() < 1e-12  # tests_O3.jl:241-249


def test_wrap_matches_direct():
    # tests_O3.jl:232-238
    p = fixture_params()
    c = O.DQMC(p, reconstruct_seeded_conf(ld))
    g2 = O.calc_greens(c.m, 2)
    assert np.abs(c.wrap_greens(g2.copy(), 2, 1) - O.calc_greens(c.m, 3)).max() < 1e-12
    assert np.abs(c.wrap_greens(g2.copy(), 2, -1) - O.calc_greens(c.m, 1)).max() < 1e-12


@pytest.mark.parametrize("chkr", [False, True])
def test_free_fermion_closed_form(chkr):
    # lambda = 0: B is phi-independent, G(l) = (I + B^M)^-1 for every slice (tests_freefermions.jl)
    p = O.Params(4, 40, safe_mult=10, hoppings=(1.0, 0.5, 0.5, 1.0), lam=0.0, chkr=chkr)
    rng = np.random.default_rng(3)
    phi = rng.uniform(-1, 1, (3, p.N, p.slices))
    c = O.DQMC(p, phi)
    c.init()
    B = c.m.slice_matrix(0, 1.0)
    G = np.linalg.inv(np.eye(p.n) + np.linalg.matrix_power(B, p.slices))
    for _ in range(2 * p.slices):
        assert np.abs(c.greens - G).max() < 1e-10
        c.propagate()
    ldet = -np.linalg.slogdet(np.eye(p.n) + np.linalg.matrix_power(B, p.slices))[1]
    c2 = O.DQMC(p, phi)
    c2.init()
    assert abs(c2.log_det - ldet) < 1e-9 * abs(ldet)


def test_gradient_trace_formula_vs_finite_differences():
    # no reference counterpart: the CPU trace formula is pinned by central differences
    p = O.Params(4, 10, hoppings=(1.0, 0.5, 0.5, 1.0), lam=1.0, r=1.0)
    rng = np.random.default_rng(8)
    phi = rng.uniform(-1, 1, (3, p.N, p.slices))
    m = O.Model(p, phi)
    g = O.action_grad_trace(m)
    gb = O.boson_action_grad(p, phi)
    h = 1e-5
    for _ in range(8):
        a, i, l = rng.integers(3), rng.integers(p.N), rng.integers(p.slices)
        pp, pm = phi.copy(), phi.copy()
        pp[a, i, l] += h
        pm[a, i, l] -= h
        fd = (O.action(m, pp) - O.action(m, pm)) / (2 * h)
        fdb = (O.calc_boson_action(p, pp) - O.calc_boson_action(p, pm)) / (2 * h)
        assert abs(g[a, i, l] - fd) < 1e-7 * max(1.0, abs(fd)), (a, i, l, g[a, i, l], fd)
        assert abs(gb[a, i, l] - fdb) < 1e-7 * max(1.0, abs(fdb))
