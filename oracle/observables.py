"""numpy restatement of the reference's fermionic observables (test oracle only).

Follows src/fermion_measurements.jl: measure_etpc! (:388-470), measure_zfpc! (:500-562),
measure_etcdc! (:592-657), measure_zfcdc! (:688-778), measure_zfccc! (:814-899) and
measure_sfdensity (:911-923), for opdim = 3 (G(mc, i, j, greens) = greens[i, j]).
Vectorised over origins r0 and displacements r; 0-based indices throughout:
  * site (x, y) -> y + L x   (sql = reshape(1:N, L, L), siteidx :7-12)
  * Green's index of flavour f (xu, yd, xd, yu = 0..3) at site s -> f N + s  (greensidx :17)
  * result arrays [y, x] (the reference's [y+1, x+1]), r = y + L x in column-major order.
Pinned by tests/test_oracle_observables.py against a literal loop restatement of the
reference on small lattices, and by the reference's own identities (zero-frequency forms at
tau = 0 equal the equal-time forms; lattice reflection symmetry for a symmetric G).
"""
import numpy as np


def _pairs(L):
    """(ri, r0) for every origin o (rows) and displacement r (columns), and the
    displaced neighbours needed by the current operator."""
    N = L * L
    o = np.arange(N)
    r = np.arange(N)
    X0, Y0 = (o // L)[:, None], (o % L)[:, None]
    x, y = (r // L)[None, :], (r % L)[None, :]

    def site(xx, yy):
        return (yy % L) + L * (xx % L)

    ri = site(X0 + x, Y0 + y)
    r0 = np.broadcast_to(site(X0, Y0), (N, N))
    nbr = {"x": (site(X0 + x + 1, Y0 + y), np.broadcast_to(site(X0 + 1, Y0), (N, N))),
           "y": (site(X0 + x, Y0 + y + 1), np.broadcast_to(site(X0, Y0 + 1), (N, N)))}
    return ri, r0, nbr


# js of measure_etpc!/measure_zfpc! (:407, :515): (xd,xu,xu,xd), (xd,xu,yu,yd),
# (yd,yu,xu,xd), (yd,yu,yu,yd) with xu, yd, xd, yu = 0, 1, 2, 3
_JS = ((2, 0, 0, 2), (2, 0, 3, 1), (1, 3, 0, 2), (1, 3, 3, 1))


def _pc_terms(G, N, ri, r0):
    ev = []
    for j1, j2, j3, j4 in _JS:
        i1, i2, i3, i4 = j1 * N + ri, j2 * N + ri, j3 * N + r0, j4 * N + r0
        ev.append(np.real(G[i1, i4] * G[i2, i3] - G[i1, i3] * G[i2, i4]))  # _pc_ev :436-449
    e1, e2, e3, e4 = ev
    return (e1 - e2 - e3 + e4).sum(axis=0), (e1 + e2 + e3 + e4).sum(axis=0)


def measure_etpc(L, G):
    """(etpc_plus, etpc_minus) as L x L [y, x] arrays."""
    N = L * L
    ri, r0, _ = _pairs(L)
    m, p = _pc_terms(G, N, ri, r0)
    return (p / N).reshape((L, L), order="F"), (m / N).reshape((L, L), order="F")


def measure_zfpc(L, Gt0s):
    N = L * L
    ri, r0, _ = _pairs(L)
    m = np.zeros(N)
    p = np.zeros(N)
    for G in Gt0s:
        dm, dp = _pc_terms(G, N, ri, r0)
        m += dm
        p += dp
    return (p / N).reshape((L, L), order="F"), (m / N).reshape((L, L), order="F")


def _cdc_terms(Gtau, G0, Gt0, G0t, N, ri, r0, et):
    """Sum over s1, s2 of ev1 - ev2 - ev3 + ev4 and ev1 + ... + ev4 (_etcdc_ev :641-653,
    _zfcdc_ev :753-766); r1 = r2 = ri, r3 = r4 = r0 and j1 = j2, j3 = j4."""
    m = np.zeros(N, dtype=np.complex128)
    p = np.zeros(N, dtype=np.complex128)
    for s1 in (0, 2):
        for s2 in (0, 2):
            ev = []
            for fa, fb in ((s1, s2), (s1, 1 + s2), (1 + s1, s2), (1 + s1, 1 + s2)):
                i1 = fa * N + ri
                i3 = fb * N + r0
                uncorr = (1.0 - Gtau[i1, i1]) * (1.0 - G0[i3, i3])
                if et:
                    c1 = (i1 == i3).astype(float) - Gtau[i3, i1]
                    c2 = Gtau[i1, i3]
                else:
                    c1 = -G0t[i3, i1]
                    c2 = Gt0[i1, i3]
                ev.append(uncorr + c1 * c2)
            e1, e2, e3, e4 = ev
            m += (e1 - e2 - e3 + e4).sum(axis=0)
            p += (e1 + e2 + e3 + e4).sum(axis=0)
    return m, p


def measure_etcdc(L, G):
    """(etcdc_plus, etcdc_minus), complex L x L [y, x]."""
    N = L * L
    ri, r0, _ = _pairs(L)
    m, p = _cdc_terms(G, G, G, None, N, ri, r0, True)
    return (p / N).reshape((L, L), order="F"), (m / N).reshape((L, L), order="F")


def measure_zfcdc(L, G, Gt0s, G0ts):
    """with greens = one equal-time G (get_Gtau = get_G0 = greens, :713-715)."""
    N = L * L
    ri, r0, _ = _pairs(L)
    m = np.zeros(N, dtype=np.complex128)
    p = np.zeros(N, dtype=np.complex128)
    for Gt0, G0t in zip(Gt0s, G0ts):
        dm, dp = _cdc_terms(G, G, Gt0, G0t, N, ri, r0, False)
        m += dm
        p += dp
    return (p / N).reshape((L, L), order="F"), (m / N).reshape((L, L), order="F")


def measure_zfccc(L, G, Gt0s, G0ts, T):
    """(zfccc_xx, zfccc_yy) with greens = one equal-time G and T = mc.l.hopping_matrix."""
    N = L * L
    ri, r0, nbr = _pairs(L)
    out = []
    for d in ("x", "y"):
        rj, r0p = nbr[d]
        lam = np.zeros(N, dtype=np.complex128)
        for Gt0, G0t in zip(Gt0s, G0ts):
            for a1 in (0, 1):
                for a2 in (0, 1):
                    for s1 in (0, 2):
                        for s2 in (0, 2):
                            f1, f2 = a1 + s1, a2 + s2
                            ais, ajs = f1 * N + ri, f1 * N + rj
                            ap0sp, ap0psp = f2 * N + r0, f2 * N + r0p
                            tij, tji = T[ais, ajs], T[ajs, ais]
                            t00p, t0p0 = T[ap0sp, ap0psp], T[ap0psp, ap0sp]
                            v = (tij * G[ajs, ais] - tji * G[ais, ajs]) * (t00p * G[ap0psp, ap0sp] -
                                                                          t0p0 * G[ap0sp, ap0psp])
                            v = v + (-tij * t00p * G0t[ap0psp, ais] * Gt0[ajs, ap0sp]
                                     + tij * t0p0 * G0t[ap0sp, ais] * Gt0[ajs, ap0psp]
                                     + tji * t00p * G0t[ap0psp, ajs] * Gt0[ais, ap0sp]
                                     - tji * t0p0 * G0t[ap0sp, ajs] * Gt0[ais, ap0psp])
                            lam += v.sum(axis=0)
        out.append((lam / -N).reshape((L, L), order="F"))
    return out[0], out[1]


def measure_sfdensity(zfccc_xx):
    """1/4 (Re Lq[ky=0, kx=1] - Re Lq[ky=1, kx=0]) with Lq = rfft(real(zfccc_xx)) (:911-923)."""
    q = np.fft.fft2(np.real(zfccc_xx))  # axis 0 = y, axis 1 = x
    return 0.25 * (np.real(q[0, 1]) - np.real(q[1, 0]))


def is_reflection_symmetric(x, tol=1e-12):
    """fermion_measurements.jl:179-201."""
    inner = x[1:, 1:]
    for t in (inner - inner[::-1, :], inner - inner[:, ::-1], x[0, 1:] - x[0, 1:][::-1], x[1:, 0] - x[1:, 0][::-1]):
        if np.max(np.real(t)) >= tol or np.max(np.imag(t)) >= tol:
            return False
    return True
