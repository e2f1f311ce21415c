"""CPU restatement of the time-displaced Green's function measurement and of the
effective -> true Green's function transform (fermion_measurements.jl:1125-1174,
1239-1576; linalg.jl:310-339, 512-567).

Test infrastructure only: imported by tests/ and by bench.py's CPU leg as the
checker, never by the product path.
"""
import numpy as np

from .dqmc_oracle import _dense_times_sparse, _lu, calc_Bchain, decompose_udt, inv, myrdiv
from scipy.linalg import lapack


def _ldiv(A, B):
    """A \\ B with an LU factorisation of A (Julia ldiv!(lu(A), B))."""
    lu, piv = _lu(A)
    x, info = lapack.zgetrs(lu, piv, np.asarray(B, dtype=np.complex128))
    if info != 0:
        raise RuntimeError("zgetrs info=%d" % info)
    return x


def effective_greens2greens(m, G):
    """G <- hop_half_inv[1] hop_half_inv[2] G hop_half[2] hop_half[1] for the checkerboard
    (fermion_measurements.jl:1125-1141); G <- eT^-1 G eT without it (:1178-1185)."""
    o = m.ops
    if m.p.chkr:
        for i in reversed(range(len(o.chkr_hop_half))):
            G = _dense_times_sparse(G, o.chkr_hop_half[i])
        for i in reversed(range(len(o.chkr_hop_half))):
            G = np.asarray(o.chkr_hop_half_inv[i] @ G)
        return G
    tmp = G @ o.hopping_matrix_exp
    return o.hopping_matrix_exp_inv @ tmp


def effective_greens2greens_ut(m, U, T):
    """The same transform on the factors of G = U D T (fermion_measurements.jl:1143-1159,
    :1187-1193)."""
    o = m.ops
    if m.p.chkr:
        for i in reversed(range(len(o.chkr_hop_half))):
            T = _dense_times_sparse(T, o.chkr_hop_half[i])
        for i in reversed(range(len(o.chkr_hop_half))):
            U = np.asarray(o.chkr_hop_half_inv[i] @ U)
        return U, T
    return o.hopping_matrix_exp_inv @ U, T @ o.hopping_matrix_exp


def inv_one_plus_udt_scalettar(U, D, T):
    """[1 + U D T]^-1 with scales split at unity and two inner UDTs (linalg.jl:310-339)."""
    Dp = np.maximum(D, 1.0)
    Dm = np.minimum(D, 1.0)
    Dpinv = 1.0 / Dp
    l = _ldiv(T, np.diag(Dpinv).astype(np.complex128))
    r = U * Dm[None, :] + l
    u, d, t = decompose_udt(r)
    r = _ldiv(t, np.diag(1.0 / d).astype(np.complex128))
    l = r @ u.conj().T
    l = Dpinv[:, None] * l
    u, d, t = decompose_udt(l)
    l = _ldiv(T, u)
    l = l * d[None, :]
    return l @ t


def inv_sum_udts_scalettar(Ua, Da, Ta, Ub, Db, Tb):
    """[Ua Da Ta + Ub Db Tb]^-1 with scales split at unity (linalg.jl:512-567)."""
    Dap = np.maximum(Da, 1.0)
    Dam = np.minimum(Da, 1.0)
    Dbp = np.maximum(Db, 1.0)
    Dbm = np.minimum(Db, 1.0)
    mat1 = myrdiv(Ta, Tb)  # Ta / Tb
    mat1 = (mat1 * Dam[:, None]) / Dbp[None, :]
    mat2 = Ua.conj().T @ Ub
    mat2 = (mat2 * Dbm[None, :]) / Dap[:, None]
    mat1 = mat1 + mat2
    U, D, T = decompose_udt(mat1)
    T = D[:, None] * T
    mat1 = _ldiv(T, U.conj().T)
    mat1 = (mat1 / Dbp[:, None]) / Dap[None, :]
    U, D, T = decompose_udt(mat1)
    mat1 = _ldiv(Tb, U)
    mat2 = T @ Ua.conj().T
    return (mat1 * D[None, :]) @ mat2


def calc_Bchain_udts(m, invert, direction):
    """UDTs at the checkpoint slices of the four B-chains (fermion_measurements.jl:1425-1507).

    direction 'left':  invert=False: B(tau, 1) = B(tau) ... B(1);  invert=True: its inverse.
    direction 'right': invert=False: B(beta, tau) = B(M) ... B(tau); invert=True: its inverse.
    Entry k (0-based) covers slices 1 .. (k+1)s ('left') or k s + 1 .. M ('right').
    Returns lists U, D, T indexed like the reference's stacks after the final reverse!.
    """
    p = m.p
    s = p.safe_mult
    nr = p.slices // s
    ranges = [range(i * s, (i + 1) * s) for i in range(nr)]  # 0-based slices
    rightmult = (direction == "right" and not invert) or (direction == "left" and invert)
    order = list(range(nr))
    if direction == "right":
        order = order[::-1]
    n = p.n
    us, ds, ts = [None] * nr, [None] * nr, [None] * nr
    for i, ridx in enumerate(order):
        if i == 0:
            cur = np.eye(n, dtype=np.complex128)
        else:
            cur = (ts[i - 1] if rightmult else us[i - 1]).copy()
        slices = ranges[ridx] if direction == "left" else reversed(ranges[ridx])
        for sl in slices:
            if not invert:
                cur = m.multiply_B_left(sl, cur) if direction == "left" else m.multiply_B_right(sl, cur)
            else:
                cur = m.multiply_B_inv_right(sl, cur) if direction == "left" else m.multiply_B_inv_left(sl, cur)
        if i != 0:
            cur = cur * ds[i - 1][None, :] if not rightmult else ds[i - 1][:, None] * cur
        U, D, T = decompose_udt(cur)
        ds[i] = D
        if not rightmult:
            us[i] = U
            ts[i] = T if i == 0 else T @ ts[i - 1]
        else:
            ts[i] = T
            us[i] = U if i == 0 else us[i - 1] @ U
    if direction == "right":
        us, ds, ts = us[::-1], ds[::-1], ts[::-1]
    return us, ds, ts


def measure_tdgfs(m):
    """Gt0[tau] and G0t[tau] for tau = 1..M (fermion_measurements.jl:1343-1407, fill_tdgf!
    :1513-1541). Returns (Gt0, G0t, stacks) with Gt0, G0t of shape (M, n, n), index tau-1,
    and stacks = {'BT0Inv', 'BBetaT', 'BT0', 'BBetaTInv'} -> (U, D, T) lists."""
    p = m.p
    M, s, n = p.slices, p.safe_mult, p.n
    if M % 2:
        raise ValueError("the reference's fill_tdgf! needs an even number of slices")
    st = {
        "BT0Inv": calc_Bchain_udts(m, True, "left"),
        "BBetaT": calc_Bchain_udts(m, False, "right"),
        "BT0": calc_Bchain_udts(m, False, "left"),
        "BBetaTInv": calc_Bchain_udts(m, True, "right"),
    }
    Gt0 = np.zeros((M, n, n), dtype=np.complex128)
    G0t = np.zeros((M, n, n), dtype=np.complex128)

    def udt(name, k):
        U, D, T = st[name]
        return U[k], D[k], T[k]

    for i in range(M // s):
        tau = i * s  # 0-based index of the 1-based slice i s + 1
        if i != 0:
            Gt0[tau] = effective_greens2greens(m, inv_sum_udts_scalettar(*udt("BT0Inv", i - 1), *udt("BBetaT", i)))
            G0t[tau] = effective_greens2greens(m, inv_sum_udts_scalettar(*udt("BT0", i - 1), *udt("BBetaTInv", i)))
        else:
            Gt0[tau] = effective_greens2greens(m, inv_one_plus_udt_scalettar(*udt("BBetaT", 0)))
            G0t[tau] = effective_greens2greens(m, inv_one_plus_udt_scalettar(*udt("BBetaTInv", 0)))
    # fill_tdgf! (1-based tau as in the reference; multiply_* take the 0-based slice)
    mhalf = M // 2 + 1
    for tau in range(mhalf, M + 1):
        if (tau - 1) % s == 0:
            continue
        Gt0[tau - 1] = m.multiply_B_left(tau - 1, Gt0[tau - 2])
        G0t[tau - 1] = m.multiply_B_inv_right(tau - 1, G0t[tau - 2])
    for tau in range(mhalf - 1, 0, -1):
        if (tau - 1) % s == 0:
            continue
        Gt0[tau - 1] = m.multiply_B_inv_left(tau, Gt0[tau])
        G0t[tau - 1] = m.multiply_B_right(tau, G0t[tau])
    G0t *= -1
    return Gt0, G0t, st


def calc_tdgf_direct(m, slice_):
    """G(slice, 1) = [B(slice-1, 1)^-1 + B(beta, slice)]^-1 from scratch
    (fermion_measurements.jl:1553-1576, scalettar=true)."""
    p = m.p
    n = p.n
    eye = np.eye(n, dtype=np.complex128)
    if slice_ != 1:
        Ul, Dl, Tl = calc_Bchain(m, 1, slice_ - 1, "inv")
    else:
        Ul, Dl, Tl = eye, np.ones(n), eye
    if slice_ != p.slices:
        Ur, Dr, Tr = calc_Bchain(m, slice_, p.slices, "B")
    else:
        Ur, Dr, Tr = eye, np.ones(n), eye
    U, D, T = _inv_sum_udts_scalettar_udt(Ul, Dl, Tl, Ur, Dr, Tr)
    U, T = effective_greens2greens_ut(m, U, T)
    return (U * D[None, :]) @ T


def _inv_sum_udts_scalettar_udt(Ua, Da, Ta, Ub, Db, Tb):
    """inv_sum_udts_scalettar in its allocating form, returning U, D, T (linalg.jl:442-496):
    the same algebra as inv_sum_udts_scalettar above with explicit inv(Tb)."""
    Dap = np.maximum(Da, 1.0)
    Dam = np.minimum(Da, 1.0)
    Dbp = np.maximum(Db, 1.0)
    Dbm = np.minimum(Db, 1.0)
    Tbinv = inv(Tb)
    mat1 = ((Ta @ Tbinv) * Dam[:, None]) / Dbp[None, :]
    mat2 = ((Ua.conj().T @ Ub) * Dbm[None, :]) / Dap[:, None]
    U, D, T = decompose_udt(mat1 + mat2)
    T = D[:, None] * T
    mat1 = (_ldiv(T, U.conj().T) / Dbp[:, None]) / Dap[None, :]
    U, D, T = decompose_udt(mat1)
    return Tbinv @ U, D, T @ Ua.conj().T
