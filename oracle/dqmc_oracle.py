"""numpy/LAPACK restatement of the reference DQMC sweep path (test oracle only).

Conventions (all 0-based here; the reference is 1-based):
  * site s = y + x*L, neighbours [up, right, down, left]      lattice.jl:105-117
  * fermion index f*N + s, flavour blocks (x, y, x, y)         hoppings.jl:84-87
  * phi has shape (3, N, M), column-major like Julia           parameters.jl:19
  * G is the n x n complex128 equal-time Green's function, n = 4N
"""
import math

import numpy as np
import scipy.sparse as sp
from scipy.linalg import expm, lapack

# ----------------------------------------------------------------------------
# Parameters (parameters.jl:4-80, 88-190)
# ----------------------------------------------------------------------------


class Params:
    """Subset of the reference Params used on the sweep path (parameters.jl:4-80)."""

    def __init__(self, L, slices, delta_tau=0.1, safe_mult=10, hoppings=(1.0, 0.5, -0.5, -1.0),
                 mu1=-0.5, mu2=None, lam=1.0, r=1.0, c=3.0, u=1.0, chkr=True, bfield=False,
                 box=0.5, all_checks=True, edrun=False, generic=False, sparsity_limit=0.01):
        self.L = int(L)
        self.slices = int(slices)
        self.delta_tau = float(delta_tau)
        self.safe_mult = int(safe_mult)
        # HOPPINGS "a,b,c,d" -> reshape(.., (2,2)): t[0,f] horizontal, t[1,f] vertical
        # (lattice.jl:59, hoppings.jl:25-43)
        h = [float(x) for x in hoppings]
        self.t = np.array([[h[0], h[2]], [h[1], h[3]]])
        self.mu1 = float(mu1)
        self.mu2 = float(mu1 if mu2 is None else mu2)
        self.lam = float(lam)
        self.r = float(r)
        self.c = float(c)
        self.u = float(u)
        self.chkr = bool(chkr)
        # DQMC(p; force_generic_chkr) (dqmc_framework.jl:105-115): CBGeneric for odd L or on request
        self.generic = self.chkr and (bool(generic) or self.L % 2 == 1)
        self.sparsity_limit = float(sparsity_limit)  # parameters.jl:70
        self.bfield = bool(bfield)
        self.box = float(box)
        self.all_checks = bool(all_checks)
        self.edrun = bool(edrun)
        self.opdim = 3
        self.flv = 4
        self.N = self.L * self.L
        self.n = self.flv * self.N
        assert self.slices % self.safe_mult == 0

    @property
    def beta(self):
        return self.slices * self.delta_tau


# ----------------------------------------------------------------------------
# Lattice (lattice.jl:105-153)
# ----------------------------------------------------------------------------


def neighbors(L):
    """4 x N table [up; right; down; left] (lattice.jl:105-117), 0-based."""
    N = L * L
    nb = np.empty((4, N), dtype=np.int64)
    for x in range(L):
        for y in range(L):
            s = y + x * L
            nb[0, s] = (y + 1) % L + x * L
            nb[1, s] = y + ((x + 1) % L) * L
            nb[2, s] = (y - 1) % L + x * L
            nb[3, s] = y + ((x - 1) % L) * L
    return nb


def time_neighbors(M):
    """2 x M table [later; earlier], periodic (lattice.jl:144-153), 0-based."""
    tn = np.empty((2, M), dtype=np.int64)
    for s in range(M):
        tn[0, s] = 0 if s == M - 1 else s + 1
        tn[1, s] = M - 1 if s == 0 else s - 1
    return tn


# ----------------------------------------------------------------------------
# Peierls phases and hopping matrices (hoppings.jl:16-254)
# ----------------------------------------------------------------------------


def peirls_phases(p):
    """phis[s][f] as N x N (trg, src) matrices (hoppings.jl:97-156), nearest neighbours only."""
    L, N = p.L, p.N
    B = np.zeros((2, 2))
    if p.bfield:
        B[0, 0] = B[1, 1] = 2 * math.pi / N
        B[0, 1] = B[1, 0] = -2 * math.pi / N
    out = [[None, None], [None, None]]
    for f in range(2):
        for s in range(2):
            phis = np.full((L, L, L, L), np.nan)  # [x, y, x', y'] 0-based
            for x in range(L):
                for y in range(L):
                    xp = (x + 1) % L
                    yp = (y + 1) % L
                    phis[x, y, x, yp] = 0.0
                    phis[x, yp, x, y] = 0.0
                    phis[x, y, xp, y] = -B[s, f] * y
                    phis[xp, y, x, y] = -phis[x, y, xp, y]
                    if y == L - 1:
                        phis[x, y, x, yp] = B[s, f] * L * x
                        phis[x, yp, x, y] = -phis[x, y, x, yp]
            # reshape(permutedims(phis, [2,1,4,3]), (N,N)) in Julia column-major:
            # element [trg, src] with trg = y + x*L, src = y' + x'*L
            per = np.transpose(phis, (1, 0, 3, 2))  # [y, x, y', x']
            out[s][f] = per.reshape((N, N), order="F")
    return out


def hopping_matrices_dense(p):
    """(hopping_matrix, exp(-dtau T/2), exp(+dtau T/2)) dense n x n.

    BFIELD=false: hoppings.jl:16-95 (flavour blocks x, y, x, y).
    BFIELD=true:  hoppings.jl:158-254 (blocks T[1,1], T[2,2], T[2,1], T[1,2]).
    """
    N = p.N
    nb = neighbors(p.L)
    if not p.bfield:
        Tx = np.diag(np.full(N, -p.mu1))
        Ty = np.diag(np.full(N, -p.mu2))
        for src in range(N):
            for k in (1, 3):  # horizontal neighbours
                trg = nb[k, src]
                Tx[trg, src] += -p.t[0, 0]
                Ty[trg, src] += -p.t[0, 1]
            for k in (0, 2):  # vertical neighbours
                trg = nb[k, src]
                Tx[trg, src] += -p.t[1, 0]
                Ty[trg, src] += -p.t[1, 1]
        blocks = [Tx, Ty, Tx, Ty]
    else:
        ph = peirls_phases(p)
        T = [[None, None], [None, None]]
        for f in range(2):
            for s in range(2):
                Tm = np.diag(np.full(N, -(p.mu1 if f == 0 else p.mu2))).astype(np.complex128)
                for src in range(N):
                    for k in (1, 3):
                        trg = nb[k, src]
                        Tm[trg, src] += -np.exp(1j * ph[s][f][trg, src]) * p.t[0, f]
                    for k in (0, 2):
                        trg = nb[k, src]
                        Tm[trg, src] += -np.exp(1j * ph[s][f][trg, src]) * p.t[1, f]
                T[s][f] = Tm
        blocks = [T[0][0], T[1][1], T[1][0], T[0][1]]
    dt = p.delta_tau
    Tfull = _blockdiag(blocks)
    eminus = _blockdiag([expm(-0.5 * dt * b) for b in blocks])
    eplus = _blockdiag([expm(0.5 * dt * b) for b in blocks])
    return Tfull, eminus, eplus


def _blockdiag(blocks):
    m = sum(b.shape[0] for b in blocks)
    dtype = np.result_type(*blocks)
    out = np.zeros((m, m), dtype=dtype)
    o = 0
    for b in blocks:
        k = b.shape[0]
        out[o:o + k, o:o + k] = b
        o += k
    return out


# ----------------------------------------------------------------------------
# Assaad 4-site checkerboard (hoppings_checkerboard.jl:1-270)
# ----------------------------------------------------------------------------


def four_site_corners(L):
    """A and B plaquette corners (hoppings_checkerboard.jl:4-18), 0-based."""
    nb = neighbors(L)
    A = [y + x * L for x in range(0, L, 2) for y in range(0, L, 2)]
    A = np.array(A, dtype=np.int64)
    B = nb[0, nb[1, A]]
    return A, B


def _plaquette_exp_real(p, corner, tflv, prefac, inverse):
    """Closed-form 4-site exponential (hoppings_checkerboard.jl:36-56): (sites, 4x4)."""
    nb = neighbors(p.L)
    i = corner
    j = nb[0, i]
    m = nb[1, i]
    n = nb[1, j]
    th = p.t[0, tflv]
    tv = p.t[1, tflv]
    fac = -prefac * p.delta_tau
    if inverse:
        fac = -fac
    cc = math.cosh(fac * -th) * math.cosh(fac * -tv)
    ss = math.sinh(fac * -th) * math.sinh(fac * -tv)
    cs = math.cosh(fac * -th) * math.sinh(fac * -tv)
    sc = math.sinh(fac * -th) * math.cosh(fac * -tv)
    sites = [i, j, m, n]
    # local order (i, j, m, n)
    E = np.array([[cc, cs, sc, ss],
                  [cs, cc, ss, sc],
                  [sc, ss, cc, cs],
                  [ss, sc, cs, cc]])
    return sites, E


def _plaquette_exp_bfield(p, corner, f, s, prefac, inverse, ph):
    """Numerical 4-site exponential with Peierls phases (hoppings_checkerboard.jl:141-197)."""
    nb = neighbors(p.L)
    cw = [corner, nb[0, corner], nb[0, nb[1, corner]], nb[1, corner]]
    sh = cw[1:] + cw[:1]
    h = p.t[0, f]
    v = p.t[1, f]
    hop = -1.0 * np.array([v, h, v, h, v, h, v, h], dtype=np.complex128)
    for k in range(4):
        i, j = cw[k], sh[k]
        hop[k] *= np.exp(1j * ph[s][f][i, j])
        hop[k + 4] *= np.exp(1j * ph[s][f][j, i])
    loc = {site: a for a, site in enumerate(cw)}
    H = np.zeros((4, 4), dtype=np.complex128)
    rows = cw + sh
    cols = sh + cw
    for k in range(8):
        H[loc[rows[k]], loc[cols[k]]] += hop[k]
    fac = -prefac * p.delta_tau
    if inverse:
        fac = -fac
    E = expm(fac * H)
    E[np.abs(E) < 1e-15] = 0  # rem_eff_zeros! (helpers.jl)
    return cw, E


def checkerboard_plaquettes(p, prefac, inverse):
    """Plaquette list for groups (A, B) x flavour blocks 0..3: list of (sites, 4x4).

    Builds what init_checkerboard_matrices(_Bfield) assembles into chkr_hop(_half)(_inv)
    (hoppings_checkerboard.jl:65-136, 199-270). Returns groups[g][b] = [(sites, E), ...].
    """
    A, Bc = four_site_corners(p.L)
    corners = [A, Bc]
    groups = []
    if not p.bfield:
        for g in range(2):
            blocks = []
            for b in range(4):
                tflv = b % 2  # x, y, x, y
                blocks.append([_plaquette_exp_real(p, c, tflv, prefac, inverse) for c in corners[g]])
            groups.append(blocks)
    else:
        ph = peirls_phases(p)
        # block order (s,f): (1,1), (2,2), (2,1), (1,2) -> 0-based (0,0),(1,1),(1,0),(0,1)
        sf = [(0, 0), (1, 1), (1, 0), (0, 1)]
        for g in range(2):
            blocks = []
            for (s, f) in sf:
                blocks.append([_plaquette_exp_bfield(p, c, f, s, prefac, inverse, ph) for c in corners[g]])
            groups.append(blocks)
    return groups


def plaquettes_to_sparse(p, blocks):
    """Assemble one group's plaquette lists into an n x n CSC matrix."""
    N, n = p.N, p.n
    rows, cols, vals = [], [], []
    for b in range(4):
        for sites, E in blocks[b]:
            for a in range(4):
                for c in range(4):
                    if E[a, c] != 0:
                        rows.append(b * N + sites[a])
                        cols.append(b * N + sites[c])
                        vals.append(E[a, c])
    dtype = np.complex128 if p.bfield else np.float64
    return sp.csc_matrix((np.array(vals, dtype=dtype), (rows, cols)), shape=(n, n))


# ----------------------------------------------------------------------------
# Generic checkerboard (hoppings_checkerboard_generic.jl:1-270)
# ----------------------------------------------------------------------------


def square_bonds(L):
    """Edges of the square lattice file as lattices/gen_square_lattice_xml.jl writes them:
    for every site its up bond, then its right bond; (src, trg, vector), 0-based."""
    nb = neighbors(L)
    out = []
    for s in range(L * L):
        out.append((s, int(nb[0, s]), (0.0, 1.0)))
        out.append((s, int(nb[1, s]), (1.0, 0.0)))
    return out


def build_checkerboard(N, bonds):
    """Greedy bond groups (build_checkerboard, hoppings_checkerboard_generic.jl:1-31): repeated
    passes over the bonds in file order, each taking every unused bond whose two sites are
    still free in this pass. Returns a list of groups, each a list of bond ids."""
    used = [False] * len(bonds)
    groups = []
    while not all(used):
        sites_used = [False] * N
        grp = []
        for bid, (src, trg, _) in enumerate(bonds):
            if used[bid] or sites_used[src] or sites_used[trg]:
                continue
            used[bid] = True
            sites_used[src] = sites_used[trg] = True
            grp.append(bid)
        groups.append(grp)
    return groups


def _rem_eff_zeros(X):
    X = np.array(X)
    X[np.abs(X) < 1e-15] = 0
    return X


def generic_group_exps(p, bonds, groups):
    """Per group g: the N x N hopping matrix T_g of its bonds for every (spin, flavour) and the
    four exponentials exp(-dtau/2 T), exp(+dtau/2 T), exp(-dtau T), exp(+dtau T)
    (init_checkerboard_matrices(_Bfield), hoppings_checkerboard_generic.jl:41-104, 199-270),
    assembled over the flavour blocks (s,f) = (1,1), (2,2), (2,1), (1,2)."""
    N = p.N
    ph = peirls_phases(p) if p.bfield else None
    out = {k: [] for k in ("half", "half_inv", "full", "full_inv")}
    for grp in groups:
        mats = {k: [] for k in out}
        for (s_, f) in ((0, 0), (1, 1), (1, 0), (0, 1)):
            T = np.zeros((N, N), dtype=np.complex128 if p.bfield else np.float64)
            for bid in grp:
                src, trg, v = bonds[bid]
                t = p.t[1, f] if v == (0.0, 1.0) else p.t[0, f]  # U: vertical, R: horizontal
                if p.bfield:
                    T[trg, src] += -np.exp(1j * ph[s_][f][trg, src]) * t
                    T[src, trg] += -np.exp(1j * ph[s_][f][src, trg]) * t
                else:
                    T[src, trg] += -t
                    T[trg, src] = T[src, trg]
            dt = p.delta_tau
            mats["half"].append(_rem_eff_zeros(expm(-0.5 * dt * T)))
            mats["half_inv"].append(_rem_eff_zeros(expm(0.5 * dt * T)))
            mats["full"].append(_rem_eff_zeros(expm(-dt * T)))
            mats["full_inv"].append(_rem_eff_zeros(expm(dt * T)))
        for k in out:
            out[k].append(sp.block_diag(mats[k], format="csc"))
    return out


def fold_chkr_grps(p, hop_half, hop_half_inv, hop_half_dagger):
    """fold_chkr_grps (hoppings_checkerboard_generic.jl:107-176): groups 2..G (1-based) merged
    into products until a product's density exceeds p.sparsity_limit."""
    n = p.n
    eye = sp.identity(n, dtype=np.complex128, format="csc")
    G = len(hop_half)
    folded_mats, folded = [], []
    cur = eye
    fstart, fstop = 1, 0  # 0-based group indices
    for i in range(1, G):
        if i != 1 and cur.count_nonzero() / float(n * n) > p.sparsity_limit:
            folded_mats.append(cur)
            folded.append(range(fstart, fstop + 1))
            fstart = fstop = i
            cur = eye
        else:
            fstop += 1
        cur = (hop_half[i] @ cur).tocsc()
    folded_mats.append(cur)
    folded.append(range(fstart, fstop + 1))
    res = {"F": folded_mats, "F_rev": [], "F_inv": [], "F_inv_rev": [], "F_dag": [], "F_dag_rev": []}
    for rng in folded:
        c_rev, c_inv, c_inv_rev, c_dag, c_dag_rev = eye, eye, eye, eye, eye
        for k in rng:
            c_rev = (c_rev @ hop_half[k]).tocsc()
            c_inv = (hop_half_inv[k] @ c_inv).tocsc()
            c_inv_rev = (c_inv_rev @ hop_half_inv[k]).tocsc()
            c_dag = (hop_half_dagger[k] @ c_dag).tocsc()
            c_dag_rev = (c_dag_rev @ hop_half_dagger[k]).tocsc()
        res["F_rev"].append(c_rev)
        res["F_inv"].append(c_inv)
        res["F_inv_rev"].append(c_inv_rev)
        res["F_dag"].append(c_dag)
        res["F_dag_rev"].append(c_dag_rev)
    return res


class Operators:
    """Host-side operators built once at setup (hoppings.jl, hoppings_checkerboard.jl)."""

    def __init__(self, p):
        self.p = p
        self.hopping_matrix, self.hopping_matrix_exp, self.hopping_matrix_exp_inv = hopping_matrices_dense(p)
        N = p.N
        muv = np.concatenate([np.full(N, p.mu1), np.full(N, p.mu2), np.full(N, p.mu1), np.full(N, p.mu2)])
        dt = p.delta_tau
        self.chkr_mu = sp.diags(np.exp(-dt * -muv)).tocsc()
        self.chkr_mu_inv = sp.diags(np.exp(dt * -muv)).tocsc()
        self.chkr_mu_half = sp.diags(np.exp(-0.5 * dt * -muv)).tocsc()
        self.chkr_mu_half_inv = sp.diags(np.exp(0.5 * dt * -muv)).tocsc()
        self.generic = p.generic
        if p.generic:
            self.bonds = square_bonds(p.L)
            self.groups = build_checkerboard(N, self.bonds)
            ex = generic_group_exps(p, self.bonds, self.groups)
            self.chkr_hop_half = ex["half"]
            self.chkr_hop_half_inv = ex["half_inv"]
            self.chkr_hop = ex["full"]
            self.chkr_hop_inv = ex["full_inv"]
            self.chkr_hop_half_dagger = [m.conj().T.tocsc() for m in self.chkr_hop_half]
            self.chkr_hop_dagger = [m.conj().T.tocsc() for m in self.chkr_hop]
            self.fold = fold_chkr_grps(p, self.chkr_hop_half, self.chkr_hop_half_inv, self.chkr_hop_half_dagger)
        elif p.chkr:
            self.plaq_half = checkerboard_plaquettes(p, 0.5, False)
            self.plaq_half_inv = checkerboard_plaquettes(p, 0.5, True)
            self.plaq = checkerboard_plaquettes(p, 1.0, False)
            self.plaq_inv = checkerboard_plaquettes(p, 1.0, True)
            self.chkr_hop_half = [plaquettes_to_sparse(p, g) for g in self.plaq_half]
            self.chkr_hop_half_inv = [plaquettes_to_sparse(p, g) for g in self.plaq_half_inv]
            self.chkr_hop = [plaquettes_to_sparse(p, g) for g in self.plaq]
            self.chkr_hop_inv = [plaquettes_to_sparse(p, g) for g in self.plaq_inv]
            self.chkr_hop_half_dagger = [m.conj().T.tocsc() for m in self.chkr_hop_half]
            self.chkr_hop_dagger = [m.conj().T.tocsc() for m in self.chkr_hop]


# ----------------------------------------------------------------------------
# Interaction exponentials (interactions.jl:35-140)
# ----------------------------------------------------------------------------


def eV_site_coeffs(p, phi_l, power):
    """C, S, R per site (interactions.jl:52-64)."""
    nrm = np.sqrt(np.sum(phi_l * phi_l, axis=0))
    sh = np.sinh(p.lam * p.delta_tau * nrm) / nrm
    C = np.cosh(p.lam * p.delta_tau * nrm)
    S = (1j * phi_l[1] - phi_l[0]) * power * sh
    R = (-phi_l[2]) * power * sh
    return C, S, R


def interaction_matrix_exp(p, phi, slice_, power):
    """Sparse eV = exp(-power dtau V(phi_slice)) (interactions.jl:35-88); slice_ 0-based."""
    N, n = p.N, p.n
    C, S, R = eV_site_coeffs(p, phi[:, :, slice_], power)
    cS = np.conj(S)
    z = np.zeros(N)
    entries = [((0, 0), C), ((0, 1), S), ((1, 0), cS), ((1, 1), C),
               ((0, 3), R), ((1, 2), -R), ((2, 1), -R), ((2, 2), C), ((2, 3), cS),
               ((3, 0), R), ((3, 2), S), ((3, 3), C)]
    rows, cols, vals = [], [], []
    idx = np.arange(N)
    for (bi, bj), v in entries:
        rows.append(bi * N + idx)
        cols.append(bj * N + idx)
        vals.append(np.asarray(v, dtype=np.complex128) + z)
    return sp.csc_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(n, n))


def interaction_matrix_exp_op(p, op, power):
    """4x4 site block (interactions.jl:102-140)."""
    nrm = math.sqrt(float(op[0] * op[0] + op[1] * op[1] + op[2] * op[2]))
    sh = power * math.sinh(p.lam * p.delta_tau * nrm) / nrm
    Cii = math.cosh(p.lam * p.delta_tau * nrm)
    Sii = (1j * op[1] - op[0]) * sh
    Rii = (-op[2]) * sh
    return np.array([[Cii, Sii, 0, Rii],
                     [np.conj(Sii), Cii, -Rii, 0],
                     [0, -Rii, Cii, np.conj(Sii)],
                     [Rii, 0, Sii, Cii]], dtype=np.complex128)


# ----------------------------------------------------------------------------
# Slice matrices (slice_matrices.jl:25-226)
# ----------------------------------------------------------------------------


class Model:
    """Model + operator state: p, operators, phi (3, N, M)."""

    def __init__(self, p, phi=None, ops=None):
        self.p = p
        self.ops = ops if ops is not None else Operators(p)
        self.nb = neighbors(p.L)
        self.tn = time_neighbors(p.slices)
        self.phi = np.zeros((3, p.N, p.slices)) if phi is None else np.array(phi, dtype=np.float64)

    # -- CBFalse: slice_matrices.jl:25-94 ------------------------------------
    def _slice_matrix_dense(self, l, power):
        eT = self.ops.hopping_matrix_exp
        eTinv = self.ops.hopping_matrix_exp_inv
        eV = interaction_matrix_exp(self.p, self.phi, l, power)
        if power > 0:
            tmp = np.asarray(eV.T.dot(eT.T).T)  # eT * eV
            return eT @ tmp
        tmp = eTinv @ eTinv
        return np.asarray(eV @ tmp)

    # -- generic dispatch ------------------------------------------------------
    def multiply_B_left(self, l, M):
        """M <- B_l M (slice_matrices.jl:45-54 CBFalse, :101-129 CBAssaad); l 0-based."""
        if not self.p.chkr:
            return self._slice_matrix_dense(l, 1.0) @ M
        o = self.ops
        eV = interaction_matrix_exp(self.p, self.phi, l, 1.0)
        if o.generic:  # slice_matrices.jl:236-270
            M = np.asarray(o.chkr_mu @ (eV @ M))
            for F in reversed(o.fold["F_rev"]):
                M = np.asarray(F @ M)
            M = np.asarray(o.chkr_hop[0] @ M)
            for F in o.fold["F"]:
                M = np.asarray(F @ M)
            return M
        M = eV @ M
        M = o.chkr_mu @ M
        M = o.chkr_hop_half[1] @ M
        M = o.chkr_hop[0] @ M
        M = o.chkr_hop_half[1] @ M
        return np.asarray(M)

    def multiply_B_right(self, l, M):
        """M <- M B_l (slice_matrices.jl:55-64, :131-153)."""
        if not self.p.chkr:
            return M @ self._slice_matrix_dense(l, 1.0)
        o = self.ops
        eV = interaction_matrix_exp(self.p, self.phi, l, 1.0)
        if o.generic:  # slice_matrices.jl:272-301
            for F in reversed(o.fold["F"]):
                M = _dense_times_sparse(M, F)
            M = _dense_times_sparse(M, o.chkr_hop[0])
            for F in o.fold["F_rev"]:
                M = _dense_times_sparse(M, F)
            M = _dense_times_sparse(M, o.chkr_mu)
            return _dense_times_sparse(M, eV)
        M = _dense_times_sparse(M, o.chkr_hop_half[1])
        M = _dense_times_sparse(M, o.chkr_hop[0])
        M = _dense_times_sparse(M, o.chkr_hop_half[1])
        M = _dense_times_sparse(M, o.chkr_mu)
        M = _dense_times_sparse(M, eV)
        return M

    def multiply_B_inv_left(self, l, M):
        """M <- B_l^-1 M (slice_matrices.jl:75-84, :155-177)."""
        if not self.p.chkr:
            return self._slice_matrix_dense(l, -1.0) @ M
        o = self.ops
        eV = interaction_matrix_exp(self.p, self.phi, l, -1.0)
        if o.generic:  # slice_matrices.jl:303-332
            for F in reversed(o.fold["F_inv_rev"]):
                M = np.asarray(F @ M)
            M = np.asarray(o.chkr_hop_inv[0] @ M)
            for F in o.fold["F_inv"]:
                M = np.asarray(F @ M)
            return np.asarray(eV @ (o.chkr_mu_inv @ M))
        M = o.chkr_hop_half_inv[1] @ M
        M = o.chkr_hop_inv[0] @ M
        M = o.chkr_hop_half_inv[1] @ M
        M = o.chkr_mu_inv @ M
        M = eV @ M
        return np.asarray(M)

    def multiply_B_inv_right(self, l, M):
        """M <- M B_l^-1 (slice_matrices.jl:65-74, :179-201)."""
        if not self.p.chkr:
            return M @ self._slice_matrix_dense(l, -1.0)
        o = self.ops
        eV = interaction_matrix_exp(self.p, self.phi, l, -1.0)
        M = _dense_times_sparse(M, eV)
        M = _dense_times_sparse(M, o.chkr_mu_inv)
        if o.generic:  # slice_matrices.jl:334-362
            for F in reversed(o.fold["F_inv"]):
                M = _dense_times_sparse(M, F)
            M = _dense_times_sparse(M, o.chkr_hop_inv[0])
            for F in o.fold["F_inv_rev"]:
                M = _dense_times_sparse(M, F)
            return M
        M = _dense_times_sparse(M, o.chkr_hop_half_inv[1])
        M = _dense_times_sparse(M, o.chkr_hop_inv[0])
        M = _dense_times_sparse(M, o.chkr_hop_half_inv[1])
        return M

    def multiply_daggered_B_left(self, l, M):
        """M <- B_l^dagger M (slice_matrices.jl:85-94, :203-226)."""
        if not self.p.chkr:
            return self._slice_matrix_dense(l, 1.0).conj().T @ M
        o = self.ops
        eV = interaction_matrix_exp(self.p, self.phi, l, 1.0)
        if o.generic:  # slice_matrices.jl:364-393
            for F in reversed(o.fold["F_dag_rev"]):
                M = np.asarray(F @ M)
            M = np.asarray(o.chkr_hop_dagger[0] @ M)
            for F in o.fold["F_dag"]:
                M = np.asarray(F @ M)
            return np.asarray(eV @ (o.chkr_mu @ M))
        M = o.chkr_hop_half_dagger[1] @ M
        M = o.chkr_hop_dagger[0] @ M
        M = o.chkr_hop_half_dagger[1] @ M
        M = o.chkr_mu @ M
        M = eV @ M
        return np.asarray(M)

    def slice_matrix(self, l, power=1.0):
        """Dense B_l or B_l^-1 (slice_matrices.jl:4-17)."""
        eye = np.eye(self.p.n, dtype=np.complex128)
        return self.multiply_B_left(l, eye) if power > 0 else self.multiply_B_inv_left(l, eye)

    # -- bosonic action (action.jl) -------------------------------------------
    def boson_action(self, phi=None):
        return calc_boson_action(self.p, self.phi if phi is None else phi)

    def boson_action_diff(self, site, slice_, new_op):
        return calc_boson_action_diff(self.p, self.phi, self.nb, self.tn, site, slice_, new_op)


def _dense_times_sparse(X, A):
    """C = X * A for CSC A (linalg.jl:45-59)."""
    return np.asarray((A.T @ X.T).T)


# ----------------------------------------------------------------------------
# Bosonic action (action.jl:1-101)
# ----------------------------------------------------------------------------


def calc_boson_action(p, phi):
    """S_B (action.jl:1-52)."""
    L, M = p.L, p.slices
    S = 0.0
    if not p.edrun:
        h = phi.reshape((3, L, L, M), order="F")
        t = h - np.roll(h, -1, axis=3)
        S += 0.5 / p.delta_tau * 1.0 / p.c ** 2 * float(np.sum(t * t))
        for ax in (1, 2):
            t = h - np.roll(h, -1, axis=ax)
            S += p.delta_tau * 0.5 * float(np.sum(t * t))
    sq = np.sum(phi * phi, axis=0)  # (N, M)
    if not p.edrun:
        S += float(np.sum(p.delta_tau * p.r / 2.0 * sq + p.delta_tau * p.u / 4.0 * sq * sq))
    else:
        S += float(np.sum(p.delta_tau * p.r / 2.0 * sq))
    return S


def calc_boson_action_diff(p, phi, nb, tn, site, slice_, new_op):
    """Delta S_B for one proposal (action.jl:57-101); site, slice_ 0-based."""
    old = phi[:, site, slice_]
    new = np.asarray(new_op, dtype=np.float64)
    diff = new - old
    old_sq = float(old @ old)
    new_sq = float(new @ new)
    sq_diff = new_sq - old_sq
    pow4_diff = new_sq * new_sq - old_sq * old_sq
    op_earlier = phi[:, site, tn[1, slice_]]
    op_later = phi[:, site, tn[0, slice_]]
    op_time = op_later + op_earlier
    op_space = np.zeros(3)
    for k in range(4):
        op_space = op_space + phi[:, nb[k, site], slice_]
    dS = 0.0
    if not p.edrun:
        dS += 1.0 / (p.delta_tau * p.c ** 2) * (sq_diff - float(op_time @ diff))
        dS += 0.5 * p.delta_tau * (4 * sq_diff - 2.0 * float(op_space @ diff))
        dS += p.delta_tau * (0.5 * p.r * sq_diff + 0.25 * p.u * pow4_diff)
    else:
        dS += p.delta_tau * (0.5 * p.r * sq_diff)
    return dS


# ----------------------------------------------------------------------------
# Dense linear algebra (linalg.jl)
# ----------------------------------------------------------------------------


def _geqp3(A):
    """LAPACK zgeqp3 with the optimal workspace, as Julia's qr!(A, Val(true)) calls it."""
    A = np.asarray(A, dtype=np.complex128)
    _, _, _, work, info = lapack.zgeqp3(A, lwork=-1)
    lwork = max(int(work[0].real), 1)
    qr, jpvt, tau, _, info = lapack.zgeqp3(A, lwork=lwork)
    if info != 0:
        raise RuntimeError("zgeqp3 info=%d" % info)
    return qr, jpvt - 1, tau


def _ungqr(qr, tau):
    _, work, info = lapack.zungqr(qr, tau, lwork=-1)
    lwork = max(int(work[0].real), 1)
    q, _, info = lapack.zungqr(qr, tau, lwork=lwork)
    if info != 0:
        raise RuntimeError("zungqr info=%d" % info)
    return q


def decompose_udt(A):
    """A = U diag(D) T via column-pivoted QR (linalg.jl:20-39). Returns U, D, T."""
    n = A.shape[1]
    qr, piv, tau = _geqp3(A)
    R = np.triu(qr[:n, :])
    invp = np.empty(n, dtype=np.int64)
    invp[piv] = np.arange(n)  # F.p[F.p] = 1:n
    D = np.abs(np.real(np.diag(R)))
    R = R / D[:, None]  # lmul!(Diagonal(1 ./ D), R)
    T = R[:, invp]
    U = _ungqr(qr, tau)
    return U, D, T


def multiply_safely(Ul, Dl, Tl, Ur, Dr, Tr):
    """linalg.jl:101-107."""
    mat = Tl @ Ur
    mat = Dl[:, None] * mat
    mat = mat * Dr[None, :]
    U, D, T = decompose_udt(mat)
    return Ul @ U, D, T @ Tr


def _lu(A):
    lu, piv, info = lapack.zgetrf(np.asarray(A, dtype=np.complex128))
    if info < 0:
        raise RuntimeError("zgetrf info=%d" % info)
    return lu, piv


def inv(A):
    """Julia inv(::Matrix) = getrf + getri."""
    lu, piv = _lu(A)
    lwork, info = lapack.zgetri_lwork(lu.shape[0])
    out, info = lapack.zgetri(lu, piv, lwork=max(int(np.real(lwork)), lu.shape[0]))
    if info != 0:
        raise RuntimeError("zgetri info=%d" % info)
    return out


def logdet_re(A):
    """Re(logdet(A)) via LU (LinearAlgebra.logdet)."""
    lu, _ = _lu(A)
    return float(np.sum(np.log(np.abs(np.diag(lu)))))


def myrdiv(a, b):
    """dest = a / b = (b' \\ a')' (linalg.jl:61), via LU of b with a conjugate-transpose solve."""
    lu, piv = _lu(b)
    x, info = lapack.zgetrs(lu, piv, np.asarray(a, dtype=np.complex128).conj().T.copy(), trans=2)
    if info != 0:
        raise RuntimeError("zgetrs info=%d" % info)
    return x.conj().T


def inv_one_plus_two_udts(Ul, Dl, Tl, Ur, Dr, Tr):
    """[1 + Ul Dl Tl (Ur Dr Tr)^dagger]^-1 in U, D, T form (linalg.jl:237-265)."""
    tmp = Tl @ Tr.conj().T
    tmp = tmp * Dr[None, :]
    tmp = Dl[:, None] * tmp
    U1, D1, T1 = decompose_udt(tmp)
    tmp3 = Ul @ U1
    tmp2 = T1 @ Ur.conj().T
    tmp = tmp3.conj().T @ inv(tmp2)
    tmp = tmp + np.diag(D1)
    u, d, t = decompose_udt(tmp)
    U = inv(t @ tmp2)
    T = (tmp3 @ u).conj().T
    return U, 1.0 / d, T


# ----------------------------------------------------------------------------
# Stack, Green's function and propagation (stack.jl)
# ----------------------------------------------------------------------------


class DQMC:
    """Reference DQMC chain state restated (stack.jl:45-118, dqmc_framework.jl:98-177)."""

    def __init__(self, p, phi, ops=None):
        self.m = Model(p, phi, ops)
        self.p = p
        n = p.n
        self.n_elements = p.slices // p.safe_mult + 1  # stack.jl:189
        self.ranges = [range(i * p.safe_mult, (i + 1) * p.safe_mult) for i in range(self.n_elements - 1)]
        self.u_stack = np.zeros((self.n_elements, n, n), dtype=np.complex128)
        self.d_stack = np.zeros((self.n_elements, n))
        self.t_stack = np.zeros((self.n_elements, n, n), dtype=np.complex128)
        eye = np.eye(n, dtype=np.complex128)
        self.Ul, self.Ur, self.Tl, self.Tr = eye.copy(), eye.copy(), eye.copy(), eye.copy()
        self.Dl, self.Dr = np.ones(n), np.ones(n)
        self.greens = np.zeros((n, n), dtype=np.complex128)
        self.log_det = 0.0
        self.current_slice = 0  # 1-based like the reference, 0..M+1
        self.direction = 1
        self.boson_action = self.m.boson_action()
        self.drifts = []  # max|G_wrapped - G_fresh| at each recompute (stack.jl:424-429, 475-480)
        self.U = self.T = None
        self.d = None

    @property
    def phi(self):
        return self.m.phi

    # stack.jl:251-272
    def build_stack(self):
        n = self.p.n
        self.u_stack[0] = np.eye(n)
        self.d_stack[0] = 1.0
        self.t_stack[0] = np.eye(n)
        for i in range(len(self.ranges)):
            self.add_slice_sequence_left(i)
        self.current_slice = self.p.slices + 1
        self.direction = -1

    # stack.jl:278-293 (idx 0-based: updates element idx+1 from idx)
    def add_slice_sequence_left(self, idx):
        curr_U = self.u_stack[idx].copy()
        for sl in self.ranges[idx]:
            curr_U = self.m.multiply_B_left(sl, curr_U)
        curr_U = curr_U * self.d_stack[idx][None, :]
        U, D, T = decompose_udt(curr_U)
        self.u_stack[idx + 1] = U
        self.d_stack[idx + 1] = D
        self.t_stack[idx + 1] = T @ self.t_stack[idx]

    # stack.jl:299-313 (updates element idx from idx+1)
    def add_slice_sequence_right(self, idx):
        curr_U = self.u_stack[idx + 1].copy()
        for sl in reversed(self.ranges[idx]):
            curr_U = self.m.multiply_daggered_B_left(sl, curr_U)
        curr_U = curr_U * self.d_stack[idx + 1][None, :]
        U, D, T = decompose_udt(curr_U)
        self.u_stack[idx] = U
        self.d_stack[idx] = D
        self.t_stack[idx] = T @ self.t_stack[idx + 1]

    # stack.jl:316-332 (slices 1-based like the reference)
    def wrap_greens(self, gf, curr_slice, direction):
        if direction == -1:
            gf = self.m.multiply_B_inv_left(curr_slice - 2, gf)
            gf = self.m.multiply_B_right(curr_slice - 2, gf)
        else:
            gf = self.m.multiply_B_left(curr_slice - 1, gf)
            gf = self.m.multiply_B_inv_right(curr_slice - 1, gf)
        return gf

    # stack.jl:338-369
    def calculate_greens(self):
        tmp = self.Tl @ self.Tr.conj().T
        tmp = tmp * self.Dr[None, :]
        tmp = self.Dl[:, None] * tmp
        U, D, T = decompose_udt(tmp)
        Ucomb = self.Ul @ U
        tmp2 = T @ self.Ur.conj().T
        tmp = myrdiv(Ucomb.conj().T, tmp2)
        tmp[np.diag_indices_from(tmp)] += D
        u, d, t = decompose_udt(tmp)
        self.T = inv(t @ tmp2)
        self.U = (Ucomb @ u).conj().T
        self.d = 1.0 / d
        self.greens = self.T @ (self.d[:, None] * self.U)

    # stack.jl:377-385
    def calculate_logdet(self):
        self.log_det = logdet_re(self.U) + float(np.sum(np.log(self.d))) + logdet_re(self.T)
        return self.log_det

    def _load_lr(self, left_idx, right_idx):
        if left_idx is not None:
            self.Ul = self.u_stack[left_idx].copy()
            self.Dl = self.d_stack[left_idx].copy()
            self.Tl = self.t_stack[left_idx].copy()
        if right_idx is not None:
            self.Ur = self.u_stack[right_idx].copy()
            self.Dr = self.d_stack[right_idx].copy()
            self.Tr = self.t_stack[right_idx].copy()

    def _reset_element(self, k):
        n = self.p.n
        self.u_stack[k] = np.eye(n)
        self.d_stack[k] = 1.0
        self.t_stack[k] = np.eye(n)

    # stack.jl:391-499
    def propagate(self):
        p = self.p
        M, s = p.slices, p.safe_mult
        if self.direction == 1:
            if self.current_slice % s == 0:
                self.current_slice += 1
                if self.current_slice == 1:
                    self._load_lr(None, 0)
                    self._reset_element(0)
                    self._load_lr(0, None)
                    self.calculate_greens()
                    self.calculate_logdet()
                elif 1 < self.current_slice <= M:
                    idx = (self.current_slice - 1) // s  # 1-based idx in the reference
                    self._load_lr(None, idx)  # u_stack[:, :, idx+1] (1-based) = element idx (0-based)
                    self.add_slice_sequence_left(idx - 1)
                    self._load_lr(idx, None)
                    g_wrapped = self.wrap_greens(self.greens.copy(), self.current_slice - 1, 1)
                    self.calculate_greens()
                    if p.all_checks:
                        self.drifts.append(float(np.max(np.abs(g_wrapped - self.greens))))
                else:
                    idx = self.n_elements - 1
                    self.add_slice_sequence_left(idx - 1)
                    self.direction = -1
                    self.current_slice = M + 1
                    self.propagate()
            else:
                self.greens = self.wrap_greens(self.greens, self.current_slice, 1)
                self.current_slice += 1
        else:
            if (self.current_slice - 1) % s == 0:
                self.current_slice -= 1
                if self.current_slice == M:
                    last = self.n_elements - 1
                    self._load_lr(last, None)
                    self._reset_element(last)
                    self._load_lr(None, last)
                    self.calculate_greens()
                    self.calculate_logdet()
                    self.greens = self.wrap_greens(self.greens, self.current_slice + 1, -1)
                elif 0 < self.current_slice < M:
                    idx = self.current_slice // s + 1  # 1-based
                    self._load_lr(idx - 1, None)
                    self.add_slice_sequence_right(idx - 1)
                    self._load_lr(None, idx - 1)
                    g_old = self.greens.copy()
                    self.calculate_greens()
                    if p.all_checks:
                        self.drifts.append(float(np.max(np.abs(g_old - self.greens))))
                    self.greens = self.wrap_greens(self.greens, self.current_slice + 1, -1)
                else:
                    self.add_slice_sequence_right(0)
                    self.direction = 1
                    self.current_slice = 0
                    self.propagate()
            else:
                self.greens = self.wrap_greens(self.greens, self.current_slice, -1)
                self.current_slice -= 1

    # dqmc_framework.jl:156-177 (with a given start conf)
    def init(self):
        self.boson_action = self.m.boson_action()
        self.build_stack()
        self.propagate()

    # ------------------------------------------------------------------------
    # Local updates (local_updates.jl:1-95)
    # ------------------------------------------------------------------------
    def calc_detratio(self, i, new_op):
        """local_updates.jl:42-59; i 0-based site. Returns det(M), sets self.delta_i, self.M."""
        p = self.p
        N = p.N
        l = self.current_slice - 1
        eVop1 = interaction_matrix_exp_op(p, self.phi[:, i, l], -1.0)
        eVop2 = interaction_matrix_exp_op(p, new_op, 1.0)
        self.delta_i = eVop1 @ eVop2 - np.eye(4)
        Mtmp = np.eye(4) - self.greens[i::N, i::N]
        self.M = np.eye(4) + self.delta_i @ Mtmp
        return complex(np.linalg.det(self.M))

    def update_greens(self, i):
        """local_updates.jl:61-95."""
        N = self.p.N
        A = self.greens[:, i::N].copy()
        for k in range(4):
            A[i + k * N, k] -= 1.0
        A = A @ inv(self.M)
        B = self.delta_i @ self.greens[i::N, :]
        self.greens = self.greens + A @ B

    def local_updates(self, stream, log=None):
        """One pass over all sites at the current slice (local_updates.jl:1-39).

        stream: object with next() -> uniform in [0,1). log: optional list that receives
        (site, accepted, p_acc, u_or_None) per proposal. Returns acceptance rate.
        """
        p = self.p
        l = self.current_slice - 1
        acc = 0
        for i in range(p.N):
            du = np.array([-p.box + 2 * p.box * stream.next() for _ in range(3)])
            new_op = self.phi[:, i, l] + du
            exp_dS = math.exp(-self.m.boson_action_diff(i, l, new_op))
            detratio = self.calc_detratio(i, new_op)
            p_acc = exp_dS * detratio.real
            if p_acc > 1.0:
                accepted, uval = True, None
            else:
                uval = stream.next()
                accepted = uval < p_acc
            if log is not None:
                log.append((i, accepted, p_acc, uval, detratio))
            if accepted:
                acc += 1
                self.phi[:, i, l] = new_op
                self.boson_action += -math.log(exp_dS)
                self.update_greens(i)
        return acc / p.N

    def update(self, stream, log=None, i=1, global_rate=0, box_global=None):
        """update(mc, i) (dqmc_framework.jl:500-517); global updates when global_rate > 0."""
        self.propagate()
        p = self.p
        if global_rate and self.current_slice == p.slices and self.direction == -1 and i % global_rate == 0:
            self.global_update(stream, p.box if box_global is None else box_global)
        return self.local_updates(stream, log)

    def sweep(self, stream, log=None):
        """One sweep = M update steps (half an up-down sweep, dqmc_framework.jl:259-262)."""
        acc = 0.0
        for _ in range(self.p.slices):
            acc += self.update(stream, log)
        return acc / self.p.slices

    # ------------------------------------------------------------------------
    # Global update (global_updates.jl:1-58)
    # ------------------------------------------------------------------------
    def global_update(self, stream, box_global):
        p = self.p
        assert self.current_slice == p.slices and self.direction == -1
        S_old = self.boson_action
        backup = (self.u_stack.copy(), self.d_stack.copy(), self.t_stack.copy(), self.greens.copy(),
                  self.log_det, self.phi.copy())
        for k in range(3):
            self.m.phi[k, :, :] += -box_global + 2 * box_global * stream.next()
        self.build_stack()
        self.propagate()
        self.boson_action = self.m.boson_action()
        p_boson = math.exp(-(self.boson_action - S_old))
        p_fermion = math.exp(backup[4] - self.log_det)
        p_acc = p_boson * p_fermion
        u = None
        if not p_acc > 1.0:
            u = stream.next()
        self.last_global = (p_acc, u)
        if p_acc > 1.0 or u < p_acc:
            return 1
        self.boson_action = S_old
        self.u_stack, self.d_stack, self.t_stack, self.greens, self.log_det, self.m.phi = backup
        return 0


# ----------------------------------------------------------------------------
# Direct Green's function (fermion_measurements.jl:964-1101)
# ----------------------------------------------------------------------------


def calc_Bchain(m, start, stop, left_op):
    """UDT of the product of left_op over slices start..stop (1-based inclusive), with a
    UDT decomposition every safe_mult slices and at the end (fermion_measurements.jl:964-1058)."""
    p = m.p
    n = p.n
    U = np.eye(n, dtype=np.complex128)
    D = np.ones(n)
    T = np.eye(n, dtype=np.complex128)
    ks = list(range(start, stop + 1))
    last = stop
    if left_op in ("inv", "dagger"):
        ks = ks[::-1]
        last = start
    for k in ks:
        if left_op == "B":
            U = m.multiply_B_left(k - 1, U)
        elif left_op == "inv":
            U = m.multiply_B_inv_left(k - 1, U)
        else:
            U = m.multiply_daggered_B_left(k - 1, U)
        if k % p.safe_mult == 0 or k == last:
            U = U * D[None, :]
            U, D, Tn = decompose_udt(U)
            T = Tn @ T
    return U, D, T


def calc_greens_and_logdet(m, slice_):
    """G(slice) and log|det G| from scratch (fermion_measurements.jl:1061-1101); slice 1-based."""
    p = m.p
    n = p.n
    Ur, Dr, Tr = calc_Bchain(m, slice_, p.slices, "dagger")
    if slice_ - 1 >= 1:
        Ul, Dl, Tl = calc_Bchain(m, 1, slice_ - 1, "B")
    else:
        Ul = np.eye(n, dtype=np.complex128)
        Dl = np.ones(n)
        Tl = np.eye(n, dtype=np.complex128)
    U, d, T = inv_one_plus_two_udts(Ul, Dl, Tl, Ur, Dr, Tr)
    ldet = logdet_re(U) + float(np.sum(np.log(d))) + logdet_re(T)
    return (U * d[None, :]) @ T, ldet


def calc_greens(m, slice_):
    return calc_greens_and_logdet(m, slice_)[0]


# ----------------------------------------------------------------------------
# Host uniform stream (the reference draws from Julia's global RNG; any host stream
# consumed in the same order reproduces the same accept/reject sequence)
# ----------------------------------------------------------------------------


class UniformStream:
    """Sequential uniforms from a pre-drawn array (same values the device consumes)."""

    def __init__(self, values):
        self.values = np.asarray(values, dtype=np.float64)
        self.pos = 0

    def next(self):
        v = float(self.values[self.pos])
        self.pos += 1
        return v


def host_uniforms(seed, count):
    """The host uniform stream used by tests and bench: numpy MT19937 doubles in [0,1)."""
    return np.random.Generator(np.random.MT19937(seed)).random(count)


# ----------------------------------------------------------------------------
# Adjoint dS/dphi (no counterpart in the reference; checked by finite differences)
# ----------------------------------------------------------------------------


def boson_action_grad(p, phi):
    """dS_B/dphi from action.jl:1-52 differentiated by hand."""
    L, M = p.L, p.slices
    h = phi.reshape((3, L, L, M), order="F")
    g = np.zeros_like(h)
    if not p.edrun:
        g += (2 * h - np.roll(h, -1, axis=3) - np.roll(h, 1, axis=3)) / (p.delta_tau * p.c ** 2)
        for ax in (1, 2):
            g += p.delta_tau * (2 * h - np.roll(h, -1, axis=ax) - np.roll(h, 1, axis=ax))
        sq = np.sum(h * h, axis=0)[None]
        g += p.delta_tau * (p.r * h + p.u * sq * h)
    else:
        g += p.delta_tau * p.r * h
    return g.reshape((3, p.N, M), order="F")


def _ev_op_grad(p, op):
    """(eVop(op,-1), [d eVop(op,+1)/d op_a for a = 0, 1, 2])."""
    op = np.asarray(op, dtype=np.float64)
    nrm = math.sqrt(float(op @ op))
    ld = p.lam * p.delta_tau
    x = ld * nrm
    sh = math.sinh(x) / nrm
    C = math.cosh(x)
    K = lambda S, R: np.array([[0, S, 0, R], [np.conj(S), 0, -R, 0], [0, -R, 0, np.conj(S)],  # noqa: E731
                               [R, 0, S, 0]], dtype=np.complex128)
    St, Rt = 1j * op[1] - op[0], -op[2]
    dsh = (ld * C - sh) / (nrm * nrm)
    grads = []
    for a in range(3):
        Sa = [-1.0, 1j, 0.0][a]
        Ra = [0.0, 0.0, -1.0][a]
        dE = ld * sh * op[a] * np.eye(4) + dsh * op[a] * K(St, Rt) + sh * K(Sa, Ra)
        grads.append(dE)
    return interaction_matrix_exp_op(p, op, -1.0), grads


def action_grad_trace(m, slices=None):
    """dS/dphi of S = S_B - log det(1 + B_M ... B_1) by the trace formula
    d log det / d phi_{a,i,l} = Re tr[(1 - G(l))_ii eV_i^-1 d_a eV_i], with G(l) the
    Green's function at slice l (stack.jl:336) from calc_greens. `slices` (0-based) limits
    the fermionic part to those slices (the others keep only dS_B/dphi); None = all."""
    p = m.p
    N = p.N
    g = boson_action_grad(p, m.phi)
    for l in (range(p.slices) if slices is None else slices):
        G = calc_greens(m, l + 1)
        for i in range(N):
            OG = np.eye(4) - G[i::N, i::N]
            Einv, dEs = _ev_op_grad(p, m.phi[:, i, l])
            for a in range(3):
                g[a, i, l] -= float(np.real(np.trace(OG @ (Einv @ dEs[a]))))
    return g


def action(m, phi=None):
    """S(phi) = S_B(phi) + log|det G| (log|det G| = -log det(1 + B_M...B_1))."""
    mm = m if phi is None else Model(m.p, phi, m.ops)
    return calc_boson_action(m.p, mm.phi) + calc_greens_and_logdet(mm, 1)[1]
