"""Model parameters and setup-time operators of the O(3) spin-fermion model.

Mirrors the reference's parameter handling (parameters.jl:4-190, XML format of
tools/xml_parameters.jl:8-41) and builds, once per chain, the hopping operators
that the reference's setup code produces (hoppings.jl:16-254,
hoppings_checkerboard.jl:1-270): the 4-site plaquette blocks of the Assaad
checkerboard, the chemical-potential diagonal, and the dense hopping
exponentials for runs without checkerboard. These are inputs to the device
library; nothing here is on the per-slice path.
"""
import math
import xml.etree.ElementTree as ET
from dataclasses import dataclass, field

import numpy as np
from scipy.linalg import expm


@dataclass
class Params:
    """Reference Params subset (parameters.jl:4-80); defaults follow Params()."""
    L: int
    slices: int
    delta_tau: float = 0.1
    safe_mult: int = 10
    hoppings: tuple = (1.0, 0.5, -0.5, -1.0)  # HOPPINGS "th_x,tv_x,th_y,tv_y"
    mu1: float = -0.5
    mu2: float = -0.5
    lam: float = 1.0
    r: float = 1.0
    c: float = 3.0
    u: float = 1.0
    checkerboard: bool = True
    generic_chkr: bool = False  # DQMC(p; force_generic_chkr=true) (dqmc_framework.jl:105-115)
    bfield: bool = False
    box: float = 0.5
    box_global: float = 0.5
    global_updates: bool = False
    global_rate: int = 5
    all_checks: bool = True
    edrun: bool = False
    seed: int = 4729339882041979125
    thermalization: int = 0
    measurements: int = 0
    extra: dict = field(default_factory=dict)

    @property
    def N(self):
        return self.L * self.L

    @property
    def n(self):
        return 4 * self.N

    @property
    def beta(self):
        return self.slices * self.delta_tau

    @property
    def generic(self):
        """CBGeneric: the checkerboard at odd L, or when forced (dqmc_framework.jl:105-115)."""
        return bool(self.checkerboard) and (bool(self.generic_chkr) or self.L % 2 == 1)

    @property
    def t(self):
        """t[0, f] horizontal, t[1, f] vertical hopping of flavour f (lattice.jl:59)."""
        h = [float(x) for x in self.hoppings]
        return np.array([[h[0], h[2]], [h[1], h[3]]])


def _parse_bool(s):
    return s.strip().lower() == "true"


def params_from_xml(path):
    """<PARAMETER name="KEY">value</PARAMETER> file -> Params (parameters.jl:88-190)."""
    root = ET.parse(path).getroot()
    d = {e.get("name"): (e.text or "").strip() for e in root.iter("PARAMETER")}
    return params_from_dict(d)


def params_from_dict(d):
    from fractions import Fraction
    dt = float(d["DELTA_TAU"])
    if "BETA" in d:
        slices = int(Fraction(d["BETA"]) / Fraction(d["DELTA_TAU"]))
    elif "T" in d:
        slices = int((1 / Fraction(d["T"])) / Fraction(d["DELTA_TAU"]))
    else:
        slices = int(d["SLICES"])
    if "MU1" in d:
        mu1 = float(d["MU1"])
        mu2 = float(d.get("MU2", d["MU1"]))
    else:
        mu1 = mu2 = float(d["MU"])
    p = Params(L=int(d["L"]), slices=slices, delta_tau=dt, safe_mult=int(d["SAFE_MULT"]),
               mu1=mu1, mu2=mu2, lam=float(d["LAMBDA"]), r=float(d["R"]), c=float(d["C"]), u=float(d["U"]))
    if "HOPPINGS" in d:
        p.hoppings = tuple(float(x) for x in d["HOPPINGS"].split(","))
    if "OPDIM" in d and int(d["OPDIM"]) != 3:
        raise ValueError("only the O(3) model (OPDIM=3) is on the device path")
    if "SEED" in d:
        p.seed = int(d["SEED"])
    if "GLOBAL_UPDATES" in d:
        p.global_updates = _parse_bool(d["GLOBAL_UPDATES"])
    if "GLOBAL_RATE" in d:
        p.global_rate = int(d["GLOBAL_RATE"])
    if "CHECKERBOARD" in d:
        p.checkerboard = _parse_bool(d["CHECKERBOARD"])
    if "BFIELD" in d:
        p.bfield = _parse_bool(d["BFIELD"])
    if "BOX_HALF_LENGTH" in d:
        p.box = p.box_global = float(d["BOX_HALF_LENGTH"])
    if "BOX_GLOBAL_HALF_LENGTH" in d:
        p.box_global = float(d["BOX_GLOBAL_HALF_LENGTH"])
    if "EDRUN" in d:
        p.edrun = _parse_bool(d["EDRUN"])
    if "WARMUP" in d:
        p.thermalization = int(d["WARMUP"]) // 2
    elif "THERMALIZATION" in d:
        p.thermalization = int(d["THERMALIZATION"])
    if "SWEEPS" in d:
        p.measurements = int(d["SWEEPS"]) // 2
    elif "MEASUREMENTS" in d:
        p.measurements = int(d["MEASUREMENTS"])
    p.extra = d
    return p


# ---------------------------------------------------------------------------
# Lattice tables
# ---------------------------------------------------------------------------

def neighbor_table(L):
    """[4, N] int32: up, right, down, left of site s = y + x L (lattice.jl:105-117)."""
    y, x = np.meshgrid(np.arange(L), np.arange(L), indexing="ij")
    site = lambda yy, xx: (yy % L) + (xx % L) * L  # noqa: E731
    tab = np.stack([site(y + 1, x), site(y, x + 1), site(y - 1, x), site(y, x - 1)])
    # column-major site order: s = y + x L
    return np.ascontiguousarray(tab.transpose(0, 2, 1).reshape(4, L * L)).astype(np.int32)


def plaquette_corners(L):
    """A corners at even (x, y) in x-major order, B = up(right(A)) (hoppings_checkerboard.jl:4-18)."""
    nb = neighbor_table(L)
    A = np.array([y + x * L for x in range(0, L, 2) for y in range(0, L, 2)], dtype=np.int64)
    return A, nb[0, nb[1, A]].astype(np.int64)


def peierls_phases(p):
    """Peierls phases phis[s][f][trg, src] (hoppings.jl:97-156), nearest neighbours."""
    L, N = p.L, p.N
    Bf = np.zeros((2, 2))
    if p.bfield:
        Bf[0, 0] = Bf[1, 1] = 2 * math.pi / N
        Bf[0, 1] = Bf[1, 0] = -2 * math.pi / N
    out = [[None, None], [None, None]]
    idx = lambda x, y: (y % L) + (x % L) * L  # noqa: E731
    for f in range(2):
        for s in range(2):
            ph = np.full((N, N), np.nan)
            for x in range(L):
                for y in range(L):
                    a, up, rt = idx(x, y), idx(x, y + 1), idx(x + 1, y)
                    ph[a, up] = 0.0  # [trg, src]
                    ph[up, a] = 0.0
                    ph[a, rt] = -Bf[s, f] * y
                    ph[rt, a] = Bf[s, f] * y
                    if y == L - 1:
                        ph[a, up] = Bf[s, f] * L * x
                        ph[up, a] = -Bf[s, f] * L * x
            out[s][f] = ph
    return out


# ---------------------------------------------------------------------------
# Operators
# ---------------------------------------------------------------------------

def _hopping_blocks(p):
    """Per flavour block b = 0..3 the N x N hopping matrix T_b (chemical potential included)."""
    N, L = p.N, p.L
    nb = neighbor_table(L)
    t = p.t
    mus = [p.mu1, p.mu2]
    if not p.bfield:
        blocks = []
        for f in (0, 1, 0, 1):
            T = np.diag(np.full(N, -mus[f]))
            for k in (1, 3):
                np.add.at(T, (nb[k], np.arange(N)), -t[0, f])
            for k in (0, 2):
                np.add.at(T, (nb[k], np.arange(N)), -t[1, f])
            blocks.append(T)
        return blocks
    ph = peierls_phases(p)
    blocks = []
    for (s, f) in ((0, 0), (1, 1), (1, 0), (0, 1)):
        T = np.diag(np.full(N, -mus[f])).astype(np.complex128)
        src = np.arange(N)
        for k, tt in ((1, t[0, f]), (3, t[0, f]), (0, t[1, f]), (2, t[1, f])):
            trg = nb[k]
            np.add.at(T, (trg, src), -np.exp(1j * ph[s][f][trg, src]) * tt)
        blocks.append(T)
    return blocks


def hopping_matrix(p):
    """mc.l.hopping_matrix (init_hopping_matrix, hoppings.jl:16-95 / :158-254): the n x n
    block-diagonal hopping matrix over the flavour blocks (x, y, x, y), complex128."""
    N = p.N
    T = np.zeros((p.n, p.n), dtype=np.complex128)
    for b, Tb in enumerate(_hopping_blocks(p)):
        T[b * N:(b + 1) * N, b * N:(b + 1) * N] = Tb
    return np.asfortranarray(T)


def _plaquette_blocks(p, prefac, sign):
    """4x4 blocks of exp(-sign * prefac * dtau * T_plaquette) for every (group, block, plaquette).

    Without field: closed form cosh/sinh products (hoppings_checkerboard.jl:36-56),
    corner order (i, up(i), right(i), up(right(i))). With field: exponential of the
    4-site hopping matrix with Peierls phases, corner order clockwise
    (hoppings_checkerboard.jl:141-197), entries below 1e-15 dropped.
    """
    L = p.L
    nb = neighbor_table(L)
    A, B = plaquette_corners(L)
    nplaq = p.N // 4
    sites = np.zeros((2, nplaq, 4), dtype=np.int32)
    mats = np.zeros((2, 4, nplaq, 4, 4), dtype=np.complex128)
    t = p.t
    fac = -prefac * p.delta_tau * sign
    if p.bfield:
        ph = peierls_phases(p)
    for g, corners in enumerate((A, B)):
        for q, c in enumerate(corners):
            if not p.bfield:
                sites[g, q] = [c, nb[0, c], nb[1, c], nb[1, nb[0, c]]]
            else:
                sites[g, q] = [c, nb[0, c], nb[0, nb[1, c]], nb[1, c]]
            for b in range(4):
                if not p.bfield:
                    f = b % 2
                    a1, a2 = fac * -t[0, f], fac * -t[1, f]
                    ch, sh, cv, sv = math.cosh(a1), math.sinh(a1), math.cosh(a2), math.sinh(a2)
                    # sites (i, j=up, m=right, n=up-right): vertical bonds i-j, m-n; horizontal i-m, j-n
                    E = np.array([[ch * cv, ch * sv, sh * cv, sh * sv],
                                  [ch * sv, ch * cv, sh * sv, sh * cv],
                                  [sh * cv, sh * sv, ch * cv, ch * sv],
                                  [sh * sv, sh * cv, ch * sv, ch * cv]])
                else:
                    s_, f = ((0, 0), (1, 1), (1, 0), (0, 1))[b]
                    cw = list(sites[g, q])
                    H = np.zeros((4, 4), dtype=np.complex128)
                    amp = [t[1, f], t[0, f], t[1, f], t[0, f]]  # v, h, v, h around the plaquette
                    for k in range(4):
                        i, j = cw[k], cw[(k + 1) % 4]
                        H[k, (k + 1) % 4] += -amp[k] * np.exp(1j * ph[s_][f][i, j])
                        H[(k + 1) % 4, k] += -amp[k] * np.exp(1j * ph[s_][f][j, i])
                    E = expm(fac * H)
                    E[np.abs(E) < 1e-15] = 0
                mats[g, b, q] = E
    return sites, mats


def square_bonds(L):
    """Edges of the square lattice file (lattices/gen_square_lattice_xml.jl): for every site
    its up bond, then its right bond; (src, trg, vertical), 0-based."""
    nb = neighbor_table(L)
    out = []
    for s in range(L * L):
        out.append((s, int(nb[0, s]), True))
        out.append((s, int(nb[1, s]), False))
    return out


def bond_groups(N, bonds):
    """build_checkerboard (hoppings_checkerboard_generic.jl:1-31): repeated passes over the
    bonds in file order; a pass takes every unused bond whose sites are still free in it."""
    used = [False] * len(bonds)
    groups = []
    while not all(used):
        free = [True] * N
        grp = []
        for q, (a, b, _) in enumerate(bonds):
            if used[q] or not free[a] or not free[b]:
                continue
            used[q] = True
            free[a] = free[b] = False
            grp.append(q)
        groups.append(grp)
    return groups


def _generic_blocks(p):
    """(group_offsets, bond_sites, bond_blocks) of the generic checkerboard: each bond's 2x2
    block of exp(-/+ prefac dtau T_g) (hoppings_checkerboard_generic.jl:41-104, 199-270) for the
    variants (half, half_inv, full, full_inv) and flavour blocks (s,f) = (1,1),(2,2),(2,1),(1,2).
    A group's bonds are disjoint, so exp(T_g) is the product of the bonds' 2x2 exponentials."""
    bonds = square_bonds(p.L)
    groups = bond_groups(p.N, bonds)
    order = [q for g in groups for q in g]
    offs = np.zeros(len(groups) + 1, dtype=np.int32)
    offs[1:] = np.cumsum([len(g) for g in groups])
    sites = np.array([[bonds[q][0], bonds[q][1]] for q in order], dtype=np.int32)
    t = p.t
    ph = peierls_phases(p) if p.bfield else None
    blocks = np.zeros((4, len(order), 4, 2, 2), dtype=np.complex128)
    for k, q in enumerate(order):
        src, trg, vert = bonds[q]
        for b, (s_, f) in enumerate(((0, 0), (1, 1), (1, 0), (0, 1))):
            amp = t[1, f] if vert else t[0, f]
            T = np.zeros((2, 2), dtype=np.complex128)  # local order (src, trg)
            T[1, 0] = -amp * (np.exp(1j * ph[s_][f][trg, src]) if p.bfield else 1.0)
            T[0, 1] = -amp * (np.exp(1j * ph[s_][f][src, trg]) if p.bfield else 1.0)
            for v, (prefac, sign) in enumerate(((0.5, 1.0), (0.5, -1.0), (1.0, 1.0), (1.0, -1.0))):
                E = expm(-sign * prefac * p.delta_tau * T)
                E[np.abs(E) < 1e-15] = 0  # rem_eff_zeros!
                blocks[v, k, b] = E
    return offs, sites, blocks


@dataclass
class OperatorSet:
    neighbors: np.ndarray
    plaq_sites: np.ndarray = None  # [2, N/4, 4] int32
    plaq: np.ndarray = None        # [4 variants, 2, 4, N/4, 4, 4] complex128
    expmu: np.ndarray = None       # [n]
    hop_exp: np.ndarray = None     # [n, n] complex (Fortran order)
    hop_exp_inv: np.ndarray = None
    group_offsets: np.ndarray = None  # generic checkerboard: [G + 1] int32
    bond_sites: np.ndarray = None     # [nbonds, 2] int32
    bond_blocks: np.ndarray = None    # [4 variants, nbonds, 4 blocks, 2, 2] complex128


def build_operators(p):
    """Setup-time operators for one chain (init_hopping_matrices, hoppings.jl:1-14)."""
    ops = OperatorSet(neighbors=neighbor_table(p.L))
    N = p.N
    if p.generic:
        ops.group_offsets, ops.bond_sites, ops.bond_blocks = _generic_blocks(p)
        muv = np.concatenate([np.full(N, p.mu1), np.full(N, p.mu2), np.full(N, p.mu1), np.full(N, p.mu2)])
        ops.expmu = np.exp(-p.delta_tau * -muv)
    elif p.checkerboard:
        variants = []
        sites = None
        for prefac, sign in ((0.5, 1.0), (0.5, -1.0), (1.0, 1.0), (1.0, -1.0)):
            sites, mats = _plaquette_blocks(p, prefac, sign)
            variants.append(mats)
        ops.plaq_sites = np.ascontiguousarray(sites, dtype=np.int32)
        ops.plaq = np.ascontiguousarray(np.stack(variants))
        muv = np.concatenate([np.full(N, p.mu1), np.full(N, p.mu2), np.full(N, p.mu1), np.full(N, p.mu2)])
        ops.expmu = np.exp(-p.delta_tau * -muv)
    else:
        blocks = _hopping_blocks(p)
        n = p.n
        em = np.zeros((n, n), dtype=np.complex128)
        ep = np.zeros((n, n), dtype=np.complex128)
        for b, T in enumerate(blocks):
            sl = slice(b * N, (b + 1) * N)
            em[sl, sl] = expm(-0.5 * p.delta_tau * T)
            ep[sl, sl] = expm(0.5 * p.delta_tau * T)
        ops.hop_exp = np.asfortranarray(em)
        ops.hop_exp_inv = np.asfortranarray(ep)
    return ops
