"""Parity of the HIP path (through the C ABI) with the CPU oracle and the reference fixtures.

Tolerances: 1e-12 max-abs against the reference's own fixtures (they reproduce at
~1e-15 on the CPU), 1e-10 max-abs for G against the oracle at config sizes (the
stated bound), drift < 1e-8, and an identical accept/reject sequence except for
proposals with |p_acc - u| < 1e-12.
"""
import math
import sys

import numpy as np
import pytest

from oracle import dqmc_oracle as O
from tests.conftest import load_fixture as ld
from tests.golden.reconstruct import reconstruct_seeded_conf

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


def fixture_params(amd_mod, checkerboard=True, bfield=True):
    return amd_mod.Params(L=4, slices=10, delta_tau=0.1, safe_mult=10, hoppings=(1.0, 0.5, -0.5, -1.0),
                          mu1=-0.5, mu2=-0.5, lam=0.5, r=2.0, c=3.0, u=1.0, checkerboard=checkerboard,
                          bfield=bfield)


def config_params(amd_mod, L, slices, checkerboard=True, lam=1.0, r=1.0, safe_mult=10):
    """The benchmark model point: hoppings 1.0,0.5,0.5,1.0, mu -0.5, c=3, u=1."""
    return amd_mod.Params(L=L, slices=slices, delta_tau=0.1, safe_mult=safe_mult, hoppings=(1.0, 0.5, 0.5, 1.0),
                          mu1=-0.5, mu2=-0.5, lam=lam, r=r, c=3.0, u=1.0, checkerboard=checkerboard)


def oracle_params(p):
    return O.Params(p.L, p.slices, delta_tau=p.delta_tau, safe_mult=p.safe_mult, hoppings=p.hoppings,
                    mu1=p.mu1, mu2=p.mu2, lam=p.lam, r=p.r, c=p.c, u=p.u, chkr=p.checkerboard, bfield=p.bfield,
                    box=p.box, all_checks=p.all_checks, edrun=p.edrun, generic=p.generic)


def rand_c(rng, n):
    return rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))


# --------------------------------------------------------------------------- primitives
def test_gemm_all_ops(amd):
    mod, DQMC = amd
    _say("gemm n=64 and n=144")
    rng = np.random.default_rng(0)
    for L in (4, 6):
        c = DQMC(config_params(mod, L, 10))
        n = c.p.n
        A, B = rand_c(rng, n), rand_c(rng, n)
        for oa in (0, 1):
            for ob in (0, 1):
                ref = (A.conj().T if oa else A) @ (B.conj().T if ob else B)
                got = c.gemm(A, B, oa, ob)
                assert np.abs(got - ref).max() < 1e-12 * np.abs(ref).max(), (L, oa, ob)
        c.close()


def test_inverse_and_logdet(amd):
    mod, DQMC = amd
    _say("gauss-jordan inverse n=64, 400")
    rng = np.random.default_rng(1)
    for L in (4, 10):
        c = DQMC(config_params(mod, L, 10))
        n = c.p.n
        A = rand_c(rng, n) + 4 * np.eye(n)
        X, ld_ = c.inv(A)
        assert np.abs(X @ A - np.eye(n)).max() < 1e-11
        assert abs(ld_ - np.linalg.slogdet(A)[1]) < 1e-9 * abs(ld_)
        c.close()


def test_udt_matches_lapack(amd):
    mod, DQMC = amd
    _say("pivoted QR UDT n=64 random + graded")
    rng = np.random.default_rng(2)
    c = DQMC(config_params(mod, 4, 10))
    n = c.p.n
    for A in (rand_c(rng, n), rand_c(rng, n) * np.logspace(0, -40, n)[None, rng.permutation(n)]):
        U, D, T = c.decompose_udt(A)
        Uo, Do, To = O.decompose_udt(A)
        assert np.abs(U.conj().T @ U - np.eye(n)).max() < 1e-13
        assert np.abs((U * D[None, :]) @ T - A).max() < 1e-13 * np.abs(A).max()
        assert np.abs(D / Do - 1).max() < 1e-10
        assert np.abs(U - Uo).max() < 1e-10 and np.abs(T - To).max() < 1e-9
    c.close()


@pytest.mark.parametrize("panel", [8, 16])
def test_udt_prepivoted_blocked_qr(amd, panel):
    """The stack's factorisation: U unitary, U diag(D) T = A column by column, T upper
    triangular after the pre-pivoting order, D descending for a column-graded matrix; both
    panel widths of the blocked QR (16 only up to n = 512)."""
    mod, DQMC = amd
    rng = np.random.default_rng(7)
    for L in (4, 10, 16):
        if panel == 16 and 4 * L * L > 512:
            continue
        _say("blocked pre-pivoted QR n=%d, %d-column panels" % (4 * L * L, panel))
        c = DQMC(config_params(mod, L, 10))
        c.set_qr_panel(panel)
        n = c.p.n
        scales = 10.0 ** rng.uniform(-30, 30, n)
        A = rand_c(rng, n) * scales[None, :]
        U, D, T = c.decompose_udt_prepivoted(A)
        assert np.abs(U.conj().T @ U - np.eye(n)).max() < 1e-12
        R = (U * D[None, :]) @ T
        err = np.linalg.norm(R - A, axis=0) / np.linalg.norm(A, axis=0)
        assert err.max() < 1e-12, err.max()
        order = np.argsort(-np.linalg.norm(A, axis=0), kind="stable")
        Tp = T[:, order]
        assert np.abs(np.tril(Tp, -1)).max() == 0.0
        assert np.allclose(np.abs(np.diag(Tp)), 1.0, atol=1e-12)
        # the diagonal of R follows the grading of the sorted columns
        d_ref = np.abs(np.diag(np.linalg.qr(A[:, order])[1]))
        assert np.max(np.abs(np.log(D / d_ref))) < 1e-8
        c.close()


def test_multiply_safely_kat(amd):
    # tests_linalg.jl:105-113 through the device UDT (32x32 fixture embedded in n=64 is not
    # equivalent, so the product is formed at n = 32 ... padded with an identity block)
    mod, DQMC = amd
    _say("multiply_safely KAT")
    u, d, t = ld("B_QR_U"), ld("B_QR_D"), ld("B_QR_T")
    mat = (t @ u) * d[:, None] * d[None, :]
    c = DQMC(config_params(mod, 4, 10))  # n = 64
    A = np.zeros((64, 64), complex)
    A[:32, :32] = mat
    A[32:, 32:] = np.diag(np.logspace(-30, -40, 32))  # below the KAT's smallest scale
    _, D, _ = c.decompose_udt(A)
    assert abs(D.max() / 5.8846316709257896e16 - 1) < 1e-8
    assert abs(np.sort(D)[32] / 1.9123571535539083e-24 - 1) < 1e-8
    c.close()


@pytest.mark.parametrize("variant", ["cbassaad_bfield", "cbassaad", "cbfalse"])
def test_slice_matrices_fixtures(amd, variant):
    # tests_O3.jl:283-331 (slice 3 of the seeded conf = afterlocal_hsfield slice 3)
    mod, DQMC = amd
    _say("slice matrices " + variant)
    chk = variant != "cbfalse"
    bf = variant == "cbassaad_bfield"
    p = fixture_params(mod, checkerboard=chk, bfield=bf)
    c = DQMC(p)
    c.hsfield = ld("afterlocal_hsfield")
    B = c.slice_matrix(3, 1.0)
    Binv = c.slice_matrix(3, -1.0)
    if variant != "cbassaad":
        suf = "" if bf else "_CBFalse"
        assert np.abs(B - ld("Bplus" + suf)).max() < 1e-12
        assert np.abs(Binv - ld("Bminus" + suf)).max() < 1e-12
    m = O.Model(oracle_params(p), ld("afterlocal_hsfield"))
    assert np.abs(B - m.slice_matrix(2, 1.0)).max() < 1e-12
    rng = np.random.default_rng(3)
    A = rand_c(rng, p.n)
    assert np.abs(c.multiply_B_right(3, A) - A @ B).max() < 1e-12
    assert np.abs(c.multiply_B_inv_right(3, A) - A @ Binv).max() < 1e-12
    assert np.abs(c.multiply_daggered_B_left(3, A) - B.conj().T @ A).max() < 1e-12
    assert np.abs(c.multiply_B_inv_left(3, c.multiply_B_left(3, A)) - A).max() < 1e-12
    c.close()


def test_slice_matrix_config_size(amd):
    mod, DQMC = amd
    _say("slice matrices L=10 checkerboard")
    p = config_params(mod, 10, 20)
    rng = np.random.default_rng(4)
    phi = rng.uniform(-1, 1, (3, p.N, p.slices))
    c = DQMC(p)
    c.hsfield = phi
    m = O.Model(oracle_params(p), phi)
    A = rand_c(rng, p.n)
    for which, fn in ((0, m.multiply_B_left), (1, m.multiply_B_right), (2, m.multiply_B_inv_left),
                      (3, m.multiply_B_inv_right), (4, m.multiply_daggered_B_left)):
        got = c._apply(7, which, A)
        assert np.abs(got - fn(6, A)).max() < 1e-11, which
    c.close()


@pytest.mark.parametrize("route", [0, 1])
def test_calculate_greens_kat(amd, route):
    # tests_O3.jl:215-230; route 0 = D-scale separation (blocked QR with 8- and with
    # 16-column panels), 1 = the reference's two-UDT route
    mod, DQMC = amd
    for panel in ((8, 16) if route == 0 else (8,)):
        _say("calculate_greens KAT route %d, %d-column panels" % (route, panel))
        _greens_kat_case(mod, DQMC, route, panel)


def _greens_kat_case(mod, DQMC, route, panel):
    p = fixture_params(mod)
    c = DQMC(p)
    c.set_greens_route(route)
    c.set_qr_panel(panel)
    conf = reconstruct_seeded_conf(ld)
    c.hsfield = conf
    G, ldet = c.calculate_greens(ld("Ul"), ld("Dl"), ld("Tl"), ld("Ur"), ld("Dr"), ld("Tr"))
    assert np.abs(G - ld("greens")).max() < 1e-12
    g_o, ld_o = O.calc_greens_and_logdet(O.Model(oracle_params(p), conf), 1)
    assert abs(ldet - ld_o) < 1e-10 * abs(ld_o)
    Gd, ldd = c.calc_greens_and_logdet(1)
    assert np.abs(Gd - ld("greens")).max() < 1e-12
    assert abs(ldd - ld_o) < 1e-10 * abs(ld_o)
    c.close()


def _oracle_greens(Ul, Dl, Tl, Ur, Dr, Tr):
    """G and log|det G| by the oracle's restatement of calculate_greens and calculate_logdet
    (stack.jl:338-385: column-pivoted UDTs, myrdiv, inv), on a bare stack object."""
    m = O.DQMC.__new__(O.DQMC)
    m.Ul, m.Dl, m.Tl, m.Ur, m.Dr, m.Tr = Ul, Dl, Tl, Ur, Dr, Tr
    m.calculate_greens()
    return m.greens, m.calculate_logdet()


@pytest.mark.parametrize("L", [12, 16])
def test_calculate_greens_large_n(amd, L):
    """calculate_greens at n = 576 and 1024, where the blocked QR keeps four rows per thread
    and the QR solve takes one column per workgroup: both device routes against the oracle on
    random UDT factors whose D spans 1e-8 .. 1e8 on each side (I + Bl Br^H then has a
    condition number far beyond a direct inverse's reach, which is what the UDT route is for)."""
    mod, DQMC = amd
    n = 4 * L * L
    _say("calculate_greens n=%d, both routes, against the oracle" % n)
    rng = np.random.default_rng(40 + L)

    def udt():
        U = np.linalg.qr(rand_c(rng, n))[0]
        D = np.sort(10.0 ** rng.uniform(-8, 8, n))[::-1].copy()
        T = np.eye(n) + 0.1 * np.triu(rand_c(rng, n), 1)
        return U, D, T

    Ul, Dl, Tl = udt()
    Ur, Dr, Tr = udt()
    G_ref, ld_ref = _oracle_greens(Ul, Dl, Tl, Ur, Dr, Tr)
    c = DQMC(config_params(mod, L, 10))
    for route in (0, 1):
        c.set_greens_route(route)
        G, ldet = c.calculate_greens(Ul, Dl, Tl, Ur, Dr, Tr)
        err = np.abs(G - G_ref).max() / np.abs(G_ref).max()
        assert err < 1e-10, (route, err)
        assert abs(ldet - ld_ref) < 1e-10 * abs(ld_ref), (route, ldet, ld_ref)
    c.close()


# --------------------------------------------------------------------------- chain
@pytest.mark.parametrize("route,panel", [(0, 8), (0, 16), (1, 8)])
def test_propagation_full_cycle_fixture(amd, route, panel):
    # tests_O3.jl:240-263: every slice of a full up-down cycle within 1e-12 of the direct G
    mod, DQMC = amd
    _say("propagation up-down cycle L=4 M=10 (bfield fixture model), route %d, %d-column panels" % (route, panel))
    p = fixture_params(mod)
    conf = reconstruct_seeded_conf(ld)
    c = DQMC(p)
    c.set_greens_route(route)
    c.set_qr_panel(panel)
    c.init(conf)
    assert (c.current_slice, c.direction) == (10, -1)
    o = O.DQMC(oracle_params(p), conf)
    o.init()
    assert np.abs(c.greens - o.greens).max() < 1e-12
    for _ in range(2 * p.slices):
        c.propagate()
        o.propagate()
        assert c.current_slice == o.current_slice and c.direction == o.direction
        assert np.abs(c.greens - o.greens).max() < 1e-12
    assert abs(c.log_det - o.log_det) < 1e-10 * abs(o.log_det)
    # propagate down to slice 1 then one more step: the `greens` fixture (tests_O3.jl:241-249)
    c2 = DQMC(p)
    c2.init(conf)
    while c2.current_slice != 1:
        c2.propagate()
    c2.propagate()
    assert np.abs(c2.greens - ld("greens")).max() < 1e-12
    c.close()
    c2.close()


def test_local_update_kats(amd):
    # tests_O3.jl:179-193
    mod, DQMC = amd
    _say("local update KATs (delta_i, M, det, afterupdate, afterlocal)")
    p = fixture_params(mod)
    conf = reconstruct_seeded_conf(ld)
    after = ld("afterlocal_hsfield")
    c = DQMC(p)
    c.init(conf)
    det, delta, M = c.calc_detratio(6, [0.488033, 0.0196912, 0.438309])
    assert abs(det - (1.000380293015979 + 1.0842021724855044e-18j)) < 1e-12
    assert np.abs(delta - ld("delta_i")).max() < 1e-12
    assert np.abs(M - ld("M")).max() < 1e-12
    c.update_greens(6)
    assert np.abs(c.greens - ld("afterupdate_greens")).max() < 1e-12
    # The sweep that followed: a uniform stream that reproduces its decisions. Accepted
    # sites propose afterlocal's value (with a 4th draw of 0 only when p_acc <= 1);
    # rejected sites propose a move the oracle rates p_acc < 0.9 and draw u = 0.95.
    box = p.box
    o = O.DQMC(oracle_params(p), conf.copy())
    o.init()
    o.calc_detratio(6, np.array([0.488033, 0.0196912, 0.438309]))
    o.update_greens(6)
    us = []
    for i in range(p.N):
        old = conf[:, i, 9]
        new = after[:, i, 9]
        if np.abs(new - old).max() > 1e-9:
            u3 = (new - old + box) / (2 * box)
            pacc = math.exp(-o.m.boson_action_diff(i, 9, new)) * o.calc_detratio(i, new).real
            us += list(u3) + ([0.0] if pacc <= 1.0 else [])
            o.phi[:, i, 9] = new
            o.update_greens(i)
        else:
            for cand in ([0.5, 0.5, 0.0], [0.5, 0.5, 1.0], [0.0, 0.5, 0.5], [1.0, 0.5, 0.5], [0.5, 0.0, 0.5]):
                trial = old + (-box + 2 * box * np.array(cand))
                pacc = math.exp(-o.m.boson_action_diff(i, 9, trial)) * o.calc_detratio(i, trial).real
                if pacc < 0.9:
                    us += cand + [0.95]
                    break
            else:
                raise AssertionError("no rejecting proposal found for site %d" % i)
    from dqmc_amd import UniformStream
    c.stream = UniformStream(values=np.array(us + [0.5] * 8))
    acc = c.local_updates(box)
    assert acc == 0.6875
    assert np.abs(c.greens - ld("afterlocal_greens")).max() < 1e-11
    assert np.abs(c.hsfield - after).max() < 1e-12
    assert abs(c.boson_action - 75.18407422927604) < 1e-10
    c.close()


def _run_pair(mod, DQMC, p, seed, nsweeps, per_slice=False, setup=None):
    rng = np.random.default_rng(seed)
    phi = rng.uniform(-1, 1, (3, p.N, p.slices))
    from dqmc_amd import UniformStream
    c = DQMC(p, uniforms=UniformStream(seed + 100))
    if setup is not None:
        setup(c)
    c.set_proposal_log(p.N * p.slices * nsweeps + 16)
    c.init(phi)
    o = O.DQMC(oracle_params(p), phi)
    o.init()
    ostream = O.UniformStream(O.host_uniforms(seed + 100, 8 * p.N * p.slices * nsweeps + 64))
    olog = []
    for _ in range(nsweeps):
        if per_slice:
            for _ in range(p.slices):
                c.propagate()
                c.local_updates()
        else:
            c.sweep(1)
        o.sweep(ostream, olog)
    return c, o, olog, phi


def _compare_logs(glog, olog):
    assert len(glog) == len(olog)
    flips = 0
    for k, (g, ob) in enumerate(zip(glog, olog)):
        o_acc, o_pacc, o_u = ob[1], ob[2], ob[3]
        if bool(g[2]) != bool(o_acc):
            gap = abs(o_pacc - (o_u if o_u is not None else 2.0))
            assert gap < 1e-12, "decision %d differs with |p_acc - u| = %g" % (k, gap)
            flips += 1
        else:
            assert abs(g[0] - o_pacc) < 1e-8 * max(1.0, abs(o_pacc)), (k, g[0], o_pacc)
    return flips


def test_seeded_chain_small(amd):
    """L=4, beta=2 (M=20) checkerboard chain: 10 sweeps, same uniform stream."""
    mod, DQMC = amd
    _say("seeded chain L=4 M=20, 10 sweeps, per-slice and fused sweeps")
    p = config_params(mod, 4, 20)
    c, o, olog, _ = _run_pair(mod, DQMC, p, 11, 10)
    flips = _compare_logs(c.proposal_log(), olog)
    assert flips == 0
    assert np.abs(c.greens - o.greens).max() < 1e-10
    assert np.abs(c.hsfield - o.phi).max() < 1e-12
    st = c.stats()
    assert st.drift_max < 1e-8 and abs(st.boson_action - o.boson_action) < 1e-9 * abs(o.boson_action)
    assert max(o.drifts) < 1e-8
    c.close()
    c2, o2, olog2, _ = _run_pair(mod, DQMC, p, 12, 2, per_slice=True)
    assert _compare_logs(c2.proposal_log(), olog2) == 0
    assert np.abs(c2.greens - o2.greens).max() < 1e-10
    c2.close()


@pytest.mark.parametrize("variant", [1, 2, 3, 4])
def test_local_update_variants(amd, variant):
    """G in global memory (1), G tiles resident in LDS (2), delayed updates (3) and windowed
    delayed updates (4): same decisions as the oracle, G within 1e-10, for several workgroup
    counts (L=6 and 8, beta=2)."""
    mod, DQMC = amd
    cases = [(6, 3), (6, 8), (6, 36)]
    if variant == 4:
        # the window without LDS-resident columns (end-of-window product on the matrix cores):
        # n = 256 at 4 workgroups, and n = 324 (L = 9, generic checkerboard) at 2 workgroups,
        # where neither the rows nor the 162 own columns fill whole 16 x 16 tiles
        cases += [(8, 4), (9, 2)]
    for L, wg in cases:
        _say("local-update variant %d, %d workgroups, L=%d M=20, 3 sweeps" % (variant, wg, L))
        p = config_params(mod, L, 20)

        def setup(c):
            c.set_groups(0, 0, wg)
            c.set_local_variant(variant)

        c, o, olog, _ = _run_pair(mod, DQMC, p, 50 + wg, 3, setup=setup)
        assert _compare_logs(c.proposal_log(), olog) == 0
        assert np.abs(c.greens - o.greens).max() < 1e-10
        assert c.stats().drift_max < 1e-8
        c.close()


def test_seeded_chain_dense(amd):
    """The dense model (L=4, no checkerboard) for 2 sweeps at beta=5."""
    mod, DQMC = amd
    _say("seeded chain L=4 M=50 dense (CBFalse), 2 sweeps")
    p = config_params(mod, 4, 50, checkerboard=False)
    c, o, olog, _ = _run_pair(mod, DQMC, p, 21, 2)
    assert _compare_logs(c.proposal_log(), olog) == 0
    assert np.abs(c.greens - o.greens).max() < 1e-10
    c.close()


def test_L10_beta20_size_parity(amd):
    """L=10, beta=20 (n=400, M=200): G after init and after one sweep vs the oracle."""
    mod, DQMC = amd
    _say("L=10 M=200 init + 1 sweep vs oracle (oracle ~1 min on the host)")
    p = config_params(mod, 10, 200)
    c, o, olog, _ = _run_pair(mod, DQMC, p, 31, 1)
    flips = _compare_logs(c.proposal_log(), olog)
    assert flips <= 1
    if flips == 0:
        assert np.abs(c.greens - o.greens).max() < 1e-10
    assert c.stats().drift_max < 1e-8
    c.close()


def test_gradient_finite_differences(amd):
    """dS/dphi vs central differences of S = S_B + log|det G| (oracle) and a CPU trace formula."""
    mod, DQMC = amd
    _say("gradient vs finite differences L=4 M=20")
    p = config_params(mod, 4, 20)
    rng = np.random.default_rng(41)
    phi = rng.uniform(-1, 1, (3, p.N, p.slices))
    c = DQMC(p)
    c.init(phi)
    g, S = c.grad()
    op = oracle_params(p)

    def action(ph):
        m = O.Model(op, ph)
        return O.calc_boson_action(op, ph) + O.calc_greens_and_logdet(m, 1)[1]

    assert abs(S - action(phi)) < 1e-9 * abs(S)
    # every component against the CPU FP64 trace formula (oracle.action_grad_trace)
    g_o = O.action_grad_trace(O.Model(op, phi))
    assert np.abs(g - g_o).max() < 1e-6 * max(1.0, np.abs(g_o).max())
    h = 1e-6
    for _ in range(6):
        a, i, l = rng.integers(3), rng.integers(p.N), rng.integers(p.slices)
        pp, pm = phi.copy(), phi.copy()
        pp[a, i, l] += h
        pm[a, i, l] -= h
        fd = (action(pp) - action(pm)) / (2 * h)
        assert abs(g[a, i, l] - fd) < 1e-6 * max(1.0, abs(fd)), (a, i, l, g[a, i, l], fd)
    c.close()


def test_batch_matches_single_chains(amd):
    mod, DQMC = amd
    from dqmc_amd import Batch, UniformStream
    _say("batch of 3 chains (two shapes) vs single chains")
    ps = [config_params(mod, 4, 20), config_params(mod, 6, 20), config_params(mod, 4, 20, lam=1.5)]
    rng = np.random.default_rng(51)
    confs = [rng.uniform(-1, 1, (3, p.N, p.slices)) for p in ps]
    b = Batch(ps, streams=[UniformStream(600 + k) for k in range(3)])
    b.init(confs)
    b.sweep(2)
    for k, p in enumerate(ps):
        c = DQMC(p, uniforms=UniformStream(600 + k))
        c.init(confs[k])
        c.sweep(2)
        assert np.abs(b.greens(k) - c.greens).max() < 1e-12
        assert np.abs(b.get_field(k) - c.hsfield).max() == 0
        c.close()
    b.close()


def test_hmc_trajectory_matches_cpu_leapfrog(amd):
    """An HMC trajectory of the gradient workload's kind (10 leapfrog steps, eps 0.02) on a
    small chain vs a CPU leapfrog driven by the oracle's trace-formula gradient."""
    mod, DQMC = amd
    _say("HMC trajectory L=4 M=20 vs CPU leapfrog")
    p = config_params(mod, 4, 20)
    rng = np.random.default_rng(61)
    phi = rng.uniform(-1, 1, (3, p.N, p.slices))
    mom = rng.standard_normal((3, p.N, p.slices))
    c = DQMC(p)
    c.init(phi)
    acc, dh = c.hmc_trajectory(mom, 10, 0.02, 0.0)
    assert acc and abs(dh) < 0.5
    op = oracle_params(p)
    m = O.Model(op, phi.copy())
    pm = mom.copy()
    H0 = O.action(m) + 0.5 * np.sum(pm * pm)
    pm -= 0.01 * O.action_grad_trace(m)
    for s in range(10):
        m.phi += 0.02 * pm
        pm -= (0.02 if s < 9 else 0.01) * O.action_grad_trace(m)
    H1 = O.action(m) + 0.5 * np.sum(pm * pm)
    assert np.abs(c.hsfield - m.phi).max() < 1e-9
    assert abs(dh - (H1 - H0)) < 1e-7
    # a rejected trajectory restores the field (u above every acceptance probability)
    phi_before = c.hsfield
    acc2, dh2 = c.hmc_trajectory(mom * 3.0, 10, 0.2, 2.0)
    if not acc2:
        assert np.array_equal(c.hsfield, phi_before)
    c.close()


def test_unsynchronised_sweeps_consume_the_stream(amd):
    """sweep(sync=False) twice, synchronize, then a synchronised sweep: the host stream advances
    by what each sweep drew, so the three sweeps see the oracle's uniforms in order."""
    mod, DQMC = amd
    from dqmc_amd import UniformStream
    _say("unsynchronised sweeps L=4 M=20, 3 sweeps")
    p = config_params(mod, 4, 20)
    phi = np.random.default_rng(71).uniform(-1, 1, (3, p.N, p.slices))
    c = DQMC(p, uniforms=UniformStream(171))
    c.set_proposal_log(3 * p.N * p.slices + 16)
    c.init(phi)
    c.sweep(1, sync=False)
    c.sweep(1, sync=False)
    c.synchronize()
    c.sweep(1)
    o = O.DQMC(oracle_params(p), phi)
    o.init()
    ostream = O.UniformStream(O.host_uniforms(171, 8 * p.N * p.slices * 3 + 64))
    olog = []
    for _ in range(3):
        o.sweep(ostream, olog)
    assert _compare_logs(c.proposal_log(), olog) == 0
    assert c.stream.consumed == ostream.pos
    assert np.abs(c.greens - o.greens).max() < 1e-10
    c.close()
