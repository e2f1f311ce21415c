"""Pins of oracle/observables.py (CPU only).

1. A literal loop restatement of the reference's measure_* functions
   (fermion_measurements.jl:388-470, 500-562, 592-657, 688-778, 814-899), 1-based like the
   reference, on random complex Green's functions at L = 2 and 3: equals the vectorised oracle.
2. The reference's own identity (tests_O3_measurements.jl:266-268): the zero-frequency
   pairing correlation of M copies of one G is M times the equal-time one; likewise for the
   charge density with G0t = G - 1 (the tau = 0 relation Gt0 - G0t = 1).
3. Lattice symmetry: for lambda = 0 without checkerboard the Green's function has the
   lattice's reflection symmetry, so every observable is reflection symmetric
   (is_reflection_symmetric, fermion_measurements.jl:179-201) and the charge densities real.
A planted mistake (a swapped flavour in js, a dropped minus sign in the current operator)
fails check 1.
"""
import numpy as np
import pytest

from dqmc_amd import model as PM
from oracle import dqmc_oracle as O
from oracle import observables as OB


def _siteidx(L, x, y):  # 1-based x, y -> 1-based site, mod1 wrapping (:7-12)
    return ((y - 1) % L + 1) + ((x - 1) % L) * L


def _G(Gm, i, j):  # 1-based access
    return Gm[i - 1, j - 1]


def _kd(a, b):
    return 1.0 if a == b else 0.0


def ref_pc(L, Gs):
    N = L * L
    Pm = np.zeros((L, L))
    Pp = np.zeros((L, L))
    xu, yd, xd, yu = 1, 2, 3, 4
    js = ((xd, xu, xu, xd), (xd, xu, yu, yd), (yd, yu, xu, xd), (yd, yu, yu, yd))
    for G in Gs:
        for x in range(L):
            for y in range(L):
                for x0 in range(1, L + 1):
                    for y0 in range(1, L + 1):
                        r1 = r2 = _siteidx(L, x0 + x, y0 + y)
                        r3 = r4 = _siteidx(L, x0, y0)
                        ev = []
                        for j in range(4):
                            i1 = (js[j][0] - 1) * N + r1
                            i2 = (js[j][1] - 1) * N + r2
                            i3 = (js[j][2] - 1) * N + r3
                            i4 = (js[j][3] - 1) * N + r4
                            ev.append(np.real(_G(G, i1, i4) * _G(G, i2, i3) - _G(G, i1, i3) * _G(G, i2, i4)))
                        Pm[y, x] += ev[0] - ev[1] - ev[2] + ev[3]
                        Pp[y, x] += ev[0] + ev[1] + ev[2] + ev[3]
    return Pp / N, Pm / N


def ref_cdc(L, G, Gt0s, G0ts, et):
    N = L * L
    Cm = np.zeros((L, L), dtype=complex)
    Cp = np.zeros((L, L), dtype=complex)
    X, Y, UP, DOWN = 1, 2, 0, 2

    def ev(Gt0, G0t, r1, r2, r3, r4, j1, j2, j3, j4):
        i1, i2, i3, i4 = ((j1 - 1) * N + r1, (j2 - 1) * N + r2, (j3 - 1) * N + r3, (j4 - 1) * N + r4)
        uncorr = (_kd(i1, i2) - _G(G, i2, i1)) * (_kd(i3, i4) - _G(G, i4, i3))
        if et:
            return uncorr + (_kd(i1, i4) - _G(G, i4, i1)) * _G(G, i2, i3)
        return uncorr + (-_G(G0t, i4, i1)) * _G(Gt0, i2, i3)

    for Gt0, G0t in zip(Gt0s, G0ts):
        for x0 in range(1, L + 1):
            for y0 in range(1, L + 1):
                for s1 in (UP, DOWN):
                    for s2 in (UP, DOWN):
                        for x in range(L):
                            for y in range(L):
                                ri = _siteidx(L, x0 + x, y0 + y)
                                r0 = _siteidx(L, x0, y0)
                                e1 = ev(Gt0, G0t, ri, ri, r0, r0, X + s1, X + s1, X + s2, X + s2)
                                e2 = ev(Gt0, G0t, ri, ri, r0, r0, X + s1, X + s1, Y + s2, Y + s2)
                                e3 = ev(Gt0, G0t, ri, ri, r0, r0, Y + s1, Y + s1, X + s2, X + s2)
                                e4 = ev(Gt0, G0t, ri, ri, r0, r0, Y + s1, Y + s1, Y + s2, Y + s2)
                                Cm[y, x] += e1 - e2 - e3 + e4
                                Cp[y, x] += e1 + e2 + e3 + e4
    return Cp / N, Cm / N


def ref_ccc(L, G, Gt0s, G0ts, T):
    N = L * L
    X, Y, UP, DOWN = 1, 2, 0, 2
    out = []
    for d in ("x", "y"):
        Lam = np.zeros((L, L), dtype=complex)
        for Gt0, G0t in zip(Gt0s, G0ts):
            for x0 in range(1, L + 1):
                for y0 in range(1, L + 1):
                    for a1 in (X, Y):
                        for a2 in (X, Y):
                            for s1 in (UP, DOWN):
                                for s2 in (UP, DOWN):
                                    for x in range(L):
                                        for y in range(L):
                                            ri = _siteidx(L, x0 + x, y0 + y)
                                            rj = _siteidx(L, x0 + x + 1, y0 + y) if d == "x" else \
                                                _siteidx(L, x0 + x, y0 + y + 1)
                                            r0 = _siteidx(L, x0, y0)
                                            r0p = _siteidx(L, x0 + 1, y0) if d == "x" else _siteidx(L, x0, y0 + 1)
                                            ais = (a1 + s1 - 1) * N + ri
                                            ajs = (a1 + s1 - 1) * N + rj
                                            ap0sp = (a2 + s2 - 1) * N + r0
                                            ap0psp = (a2 + s2 - 1) * N + r0p
                                            tij, tji = _G(T, ais, ajs), _G(T, ajs, ais)
                                            t00p, t0p0 = _G(T, ap0sp, ap0psp), _G(T, ap0psp, ap0sp)
                                            Lam[y, x] += ((tij * _G(G, ajs, ais) - tji * _G(G, ais, ajs)) *
                                                          (t00p * _G(G, ap0psp, ap0sp) - t0p0 * _G(G, ap0sp, ap0psp)))
                                            Lam[y, x] += (-tij * t00p * _G(G0t, ap0psp, ais) * _G(Gt0, ajs, ap0sp)
                                                          + tij * t0p0 * _G(G0t, ap0sp, ais) * _G(Gt0, ajs, ap0psp)
                                                          + tji * t00p * _G(G0t, ap0psp, ajs) * _G(Gt0, ais, ap0sp)
                                                          - tji * t0p0 * _G(G0t, ap0sp, ajs) * _G(Gt0, ais, ap0psp))
        out.append(Lam / -N)
    return out


def _rand(rng, n):
    return rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))


@pytest.mark.parametrize("L", [2, 3])
def test_vectorised_oracle_equals_literal_restatement(L):
    rng = np.random.default_rng(L)
    n = 4 * L * L
    G = _rand(rng, n)
    Gt0s = [_rand(rng, n) for _ in range(3)]
    G0ts = [_rand(rng, n) for _ in range(3)]
    T = _rand(rng, n)
    for a, b in zip(OB.measure_etpc(L, G), ref_pc(L, [G])):
        assert np.abs(a - b).max() < 1e-12
    for a, b in zip(OB.measure_zfpc(L, Gt0s), ref_pc(L, Gt0s)):
        assert np.abs(a - b).max() < 1e-12
    for a, b in zip(OB.measure_etcdc(L, G), ref_cdc(L, G, [G], [None], True)):
        assert np.abs(a - b).max() < 1e-12
    for a, b in zip(OB.measure_zfcdc(L, G, Gt0s, G0ts), ref_cdc(L, G, Gt0s, G0ts, False)):
        assert np.abs(a - b).max() < 1e-12
    for a, b in zip(OB.measure_zfccc(L, G, Gt0s, G0ts, T), ref_ccc(L, G, Gt0s, G0ts, T)):
        assert np.abs(a - b).max() < 1e-11 * max(1.0, np.abs(b).max())


def test_zero_frequency_forms_at_tau_zero():
    # tests_O3_measurements.jl:266-268 and the tau = 0 relation Gt0 - G0t = 1
    L, M = 3, 4
    rng = np.random.default_rng(5)
    n = 4 * L * L
    G = _rand(rng, n)
    ep, em = OB.measure_etpc(L, G)
    zp, zm = OB.measure_zfpc(L, [G] * M)
    assert np.abs(zp / M - ep).max() < 1e-12 and np.abs(zm / M - em).max() < 1e-12
    cp, cm = OB.measure_etcdc(L, G)
    zp, zm = OB.measure_zfcdc(L, G, [G] * M, [G - np.eye(n)] * M)
    assert np.abs(zp / M - cp).max() < 1e-11 and np.abs(zm / M - cm).max() < 1e-11


def test_free_fermion_observables_are_reflection_symmetric():
    # lambda = 0, no checkerboard: G = (1 + e^{-beta T})^-1 has the lattice symmetry
    p = PM.Params(L=4, slices=20, hoppings=(1.0, 0.5, 0.5, 1.0), lam=0.0, checkerboard=False)
    T = PM.hopping_matrix(p)
    e, V = np.linalg.eigh(T)
    G = (V * (1.0 / (1.0 + np.exp(-p.beta * e)))) @ V.conj().T
    dt = p.delta_tau
    Gt0s = [(V * (np.exp(-k * dt * e) / (1.0 + np.exp(-p.beta * e)))) @ V.conj().T for k in range(p.slices)]
    G0ts = [-(V * (np.exp(-(p.slices - k) * dt * e) / (1.0 + np.exp(-p.beta * e)))) @ V.conj().T
            for k in range(p.slices)]
    for arr in (*OB.measure_etpc(4, G), *OB.measure_zfpc(4, Gt0s), *OB.measure_etcdc(4, G),
                *OB.measure_zfcdc(4, G, Gt0s, G0ts), *OB.measure_zfccc(4, G, Gt0s, G0ts, T)):
        assert OB.is_reflection_symmetric(arr, tol=1e-10)
        assert np.abs(np.imag(arr)).max() < 1e-12
    # a mistake the symmetry check catches: a displacement bin off by one
    p_, _ = OB.measure_etpc(4, G)
    assert not OB.is_reflection_symmetric(np.roll(p_, 1, axis=1) + 1e-3 * np.arange(16).reshape(4, 4))
    assert O is not None
