"""Host-side mirror of the reference's DQMC interface over the HIP library.

Names follow the reference (dqmc_framework.jl, stack.jl, local_updates.jl,
slice_matrices.jl): `DQMC(p)`, `init(conf)` (= init!), `build_stack`,
`propagate`, `local_updates`, `wrap_greens`, `calculate_greens`,
`calc_greens_and_logdet`, `calc_detratio`, `update_greens`. Every call goes
through the C ABI of libdqmc_hip.so; nothing here computes on the host beyond
setup-time operator construction and the host uniform stream.

Slices are 1-based wherever the reference uses `current_slice` (0..M+1);
sites are 0-based.
"""
import ctypes
import sys

import numpy as np

from . import _capi
from ._capi import check, dptr, iptr
from .model import Params, build_operators


class UniformStream:
    """The host uniform stream the reference draws with rand() (local_updates.jl:9,31).

    numpy MT19937 doubles in [0, 1); `take(count)` returns the next `count` values
    without consuming them, `advance(k)` consumes k. Any stream consumed in the same
    order reproduces the same accept/reject sequence.
    """

    def __init__(self, seed=None, values=None):
        self._gen = None if values is not None else np.random.Generator(np.random.MT19937(seed))
        self._buf = np.zeros(0) if values is None else np.asarray(values, dtype=np.float64)
        self.pos = 0  # absolute position of _buf[0]
        self.consumed = 0

    def take(self, count):
        off = self.consumed - self.pos
        need = off + count - len(self._buf)
        if need > 0:
            if self._gen is None:
                raise ValueError("uniform stream exhausted")
            self._buf = np.concatenate([self._buf, self._gen.random(max(need, 1 << 16))])
        return self._buf[off:off + count]

    def advance(self, k):
        self.consumed += k
        drop = self.consumed - self.pos
        if drop > (1 << 20):
            self._buf = self._buf[drop:]
            self.pos += drop


def _operators_struct(ops, keep):
    s = _capi.Operators()
    s.neighbors = iptr(ops.neighbors)
    keep.append(ops.neighbors)
    if ops.plaq is not None:
        ps = np.ascontiguousarray(ops.plaq_sites, dtype=np.int32)
        pl = np.ascontiguousarray(ops.plaq, dtype=np.complex128)
        em = np.ascontiguousarray(ops.expmu, dtype=np.float64)
        keep += [ps, pl, em]
        s.plaq_sites = iptr(ps)
        s.plaq = pl.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        s.expmu = dptr(em)
    if ops.group_offsets is not None:
        go = np.ascontiguousarray(ops.group_offsets, dtype=np.int32)
        bs = np.ascontiguousarray(ops.bond_sites, dtype=np.int32)
        bb = np.ascontiguousarray(ops.bond_blocks, dtype=np.complex128)
        em = np.ascontiguousarray(ops.expmu, dtype=np.float64)
        keep += [go, bs, bb, em]
        s.n_groups = len(go) - 1
        s.group_offsets = iptr(go)
        s.bond_sites = iptr(bs)
        s.bond_blocks = bb.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        s.expmu = dptr(em)
    if ops.hop_exp is not None:
        he = np.asfortranarray(ops.hop_exp, dtype=np.complex128)
        hi = np.asfortranarray(ops.hop_exp_inv, dtype=np.complex128)
        keep += [he, hi]
        s.hop_exp = he.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        s.hop_exp_inv = hi.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    return s


def _params_struct(p):
    return _capi.Params(L=p.L, slices=p.slices, safe_mult=p.safe_mult,
                        checkerboard=(2 if p.generic else 1) if p.checkerboard else 0,
                        edrun=int(p.edrun), all_checks=int(p.all_checks), delta_tau=p.delta_tau, lam=p.lam,
                        r=p.r, c=p.c, u=p.u)


def _cmat_ptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _as_fcomplex(a, n):
    a = np.asfortranarray(a, dtype=np.complex128)
    if a.shape != (n, n):
        raise ValueError("expected a %dx%d matrix" % (n, n))
    return a


class DQMC:
    """One Markov chain on the device (DQMC(p) + initialize_stack, dqmc_framework.jl:105-135)."""

    def __init__(self, p: Params, device=0, uniforms=None, ops=None):
        self.p = p
        self.lib = _capi.load()
        self._keep = []
        ops = ops if ops is not None else build_operators(p)
        self._ops_struct = _operators_struct(ops, self._keep)
        self._params_struct = _params_struct(p)
        h = ctypes.c_void_p()
        check(self.lib.dqmc_ctx_create(ctypes.byref(self._params_struct), ctypes.byref(self._ops_struct),
                                       device, ctypes.byref(h)))
        self.h = h
        self.stream = uniforms if uniforms is not None else UniformStream(p.seed)
        self._log_cap = 0
        self._pending_uniforms = False  # a sweep(sync=False) whose consumption is not yet counted

    def close(self):
        if getattr(self, "h", None):
            self.lib.dqmc_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        # at interpreter shutdown the HIP runtime may already be finalised; the process
        # exit releases the device memory then
        if sys.is_finalizing():
            return
        try:
            self.close()
        except Exception:
            pass

    # -- field ----------------------------------------------------------------
    @property
    def hsfield(self):
        phi = np.zeros((3, self.p.N, self.p.slices), order="F")
        check(self.lib.dqmc_get_field(self.h, dptr(phi)))
        return phi

    @hsfield.setter
    def hsfield(self, phi):
        phi = np.asfortranarray(phi, dtype=np.float64)
        if phi.shape != (3, self.p.N, self.p.slices):
            raise ValueError("hsfield must have shape (3, N, M)")
        check(self.lib.dqmc_set_field(self.h, dptr(phi)))

    # -- init!, stack, propagation --------------------------------------------
    def init(self, conf):
        """init!(mc, conf) (dqmc_framework.jl:156-177)."""
        self.hsfield = conf
        check(self.lib.dqmc_init(self.h))

    def build_stack(self):
        check(self.lib.dqmc_build_stack(self.h))

    def propagate(self, want_drift=False):
        """propagate (stack.jl:391-499). Returns (current_slice, direction, drift or None)."""
        cs, d = ctypes.c_int(), ctypes.c_int()
        dr = ctypes.c_double(-1.0)
        check(self.lib.dqmc_propagate(self.h, ctypes.byref(cs), ctypes.byref(d),
                                      ctypes.byref(dr) if want_drift else None))
        drift = dr.value if (want_drift and dr.value >= 0) else None
        return cs.value, d.value, drift

    def _upload_uniforms(self, count):
        # uniforms of an unsynchronised sweep must be consumed before the next upload
        # (uploading resets the device cursor), else they would be handed out again
        self._settle_uniforms()
        vals = np.ascontiguousarray(self.stream.take(count))
        check(self.lib.dqmc_set_uniforms(self.h, dptr(vals), len(vals)))

    def _consume_uniforms(self):
        st = self.stats()
        self.stream.advance(st.uniforms_used)
        self._pending_uniforms = False

    def _settle_uniforms(self):
        if self._pending_uniforms:
            check(self.lib.dqmc_synchronize(self.h))
            self._consume_uniforms()

    def local_updates(self, box=None):
        """local_updates (local_updates.jl:1-39); returns the acceptance rate."""
        box = self.p.box if box is None else box
        self._upload_uniforms(4 * self.p.N)
        acc = ctypes.c_double()
        check(self.lib.dqmc_local_updates(self.h, box, ctypes.byref(acc)))
        self._consume_uniforms()
        return acc.value

    def global_update(self, box_global=None):
        """global_update (global_updates.jl:19-58) at current_slice == M, direction == -1;
        returns 1 if accepted, else 0 (field, stack, G, log_det and boson action restored)."""
        b = self.p.box_global if box_global is None else box_global
        self._upload_uniforms(4)
        acc = ctypes.c_int()
        check(self.lib.dqmc_global_update(self.h, b, ctypes.byref(acc)))
        self._consume_uniforms()
        return acc.value

    def update(self, i=1, box=None):
        """update(mc, i) (dqmc_framework.jl:500-517): propagate, a global update when
        p.global_updates and current_slice == M, direction == -1 and i % global_rate == 0,
        then local updates. i is the ud-sweep counter of thermalize!/measure!."""
        cs, d, _ = self.propagate()
        if self.p.global_updates and cs == self.p.slices and d == -1 and i % self.p.global_rate == 0:
            self.global_update()
        return self.local_updates(box)

    def sweep(self, nsweeps=1, box=None, sync=True):
        """nsweeps x M update steps without host round trips between slices."""
        box = self.p.box if box is None else box
        self._upload_uniforms(4 * self.p.N * self.p.slices * nsweeps)
        check(self.lib.dqmc_sweep(self.h, nsweeps, box))
        self._pending_uniforms = True
        if sync:
            self.synchronize()

    def synchronize(self):
        """Waits for the chain's queued work; the uniforms an unsynchronised sweep drew are
        then consumed from the host stream."""
        check(self.lib.dqmc_synchronize(self.h))
        if self._pending_uniforms:
            self._consume_uniforms()

    def set_groups(self, wg_qr=0, wg_inv=0, wg_local=0):
        check(self.lib.dqmc_ctx_set_groups(self.h, wg_qr, wg_inv, wg_local))

    def set_udt_route(self, route):
        """0: blocked QR with pre-pivoting for the stack (default); 1: column-pivoted QR."""
        check(self.lib.dqmc_set_udt_route(self.h, route))

    def set_local_variant(self, variant):
        """0 auto (default); 1 G in global memory; 2 LDS-resident G tiles; 3 delayed updates;
        4 windowed delayed updates."""
        check(self.lib.dqmc_set_local_variant(self.h, variant))

    def set_qr_panel(self, nb):
        """Panel width of the blocked QR: 16 (n <= 512) or 8; 0 restores the default."""
        check(self.lib.dqmc_set_qr_panel(self.h, nb))

    def use_graph(self, enable=True):
        """Capture each sweep's kernel sequence in a HIP graph (one per starting state) and
        replay it; same results as eager launches."""
        check(self.lib.dqmc_use_graph(self.h, 1 if enable else 0))

    def autotune(self, reps=3, box=None):
        """Time the local updates over (variant, workgroups) and the blocked QR over
        (workgroups, panel width) on the current state and keep the fastest; the state is
        restored exactly. Returns the choice as a dict."""
        box = self.p.box if box is None else box
        rep = np.zeros(6)
        check(self.lib.dqmc_autotune(self.h, box, reps, dptr(rep)))
        return _tune_report(rep)

    def set_greens_route(self, route):
        """0: D-scale separation (default); 1: the reference's calculate_greens verbatim."""
        check(self.lib.dqmc_set_greens_route(self.h, route))

    # -- state ---------------------------------------------------------------
    @property
    def greens(self):
        G = np.zeros((self.p.n, self.p.n), dtype=np.complex128, order="F")
        check(self.lib.dqmc_get_greens(self.h, _cmat_ptr(G)))
        return G

    @greens.setter
    def greens(self, G):
        G = _as_fcomplex(G, self.p.n)
        check(self.lib.dqmc_set_greens(self.h, _cmat_ptr(G)))

    def stack_element(self, k):
        n = self.p.n
        U = np.zeros((n, n), dtype=np.complex128, order="F")
        T = np.zeros((n, n), dtype=np.complex128, order="F")
        D = np.zeros(n)
        check(self.lib.dqmc_get_stack(self.h, k, _cmat_ptr(U), dptr(D), _cmat_ptr(T)))
        return U, D, T

    def stats(self):
        st = _capi.Stats()
        check(self.lib.dqmc_get_stats(self.h, ctypes.byref(st)))
        return st

    def reset_stats(self):
        check(self.lib.dqmc_reset_stats(self.h))

    @property
    def current_slice(self):
        return self.stats().current_slice

    @property
    def direction(self):
        return self.stats().direction

    @property
    def boson_action(self):
        return self.stats().boson_action

    @property
    def log_det(self):
        return self.stats().log_det

    def set_proposal_log(self, capacity):
        self._log_cap = capacity
        check(self.lib.dqmc_set_proposal_log(self.h, capacity))

    def proposal_log(self):
        cnt = ctypes.c_longlong()
        out = np.zeros((max(self._log_cap, 1), 4))
        check(self.lib.dqmc_get_proposal_log(self.h, dptr(out), ctypes.byref(cnt)))
        return out[:min(cnt.value, self._log_cap)]

    # -- unit surfaces --------------------------------------------------------
    def _apply(self, slice_, which, M):
        X = _as_fcomplex(M, self.p.n).copy(order="F")
        check(self.lib.dqmc_apply_B(self.h, slice_ - 1, which, _cmat_ptr(X)))
        return X

    def multiply_B_left(self, slice_, M):
        return self._apply(slice_, 0, M)

    def multiply_B_right(self, slice_, M):
        return self._apply(slice_, 1, M)

    def multiply_B_inv_left(self, slice_, M):
        return self._apply(slice_, 2, M)

    def multiply_B_inv_right(self, slice_, M):
        return self._apply(slice_, 3, M)

    def multiply_daggered_B_left(self, slice_, M):
        return self._apply(slice_, 4, M)

    def slice_matrix(self, slice_, power=1.0):
        eye = np.eye(self.p.n, dtype=np.complex128)
        return self.multiply_B_left(slice_, eye) if power > 0 else self.multiply_B_inv_left(slice_, eye)

    def wrap_greens(self, G, slice_, direction):
        X = _as_fcomplex(G, self.p.n).copy(order="F")
        check(self.lib.dqmc_wrap_greens(self.h, _cmat_ptr(X), slice_, direction))
        return X

    def decompose_udt(self, A):
        n = self.p.n
        A = _as_fcomplex(A, n)
        U = np.zeros((n, n), dtype=np.complex128, order="F")
        T = np.zeros((n, n), dtype=np.complex128, order="F")
        D = np.zeros(n)
        check(self.lib.dqmc_udt(self.h, _cmat_ptr(A), _cmat_ptr(U), dptr(D), _cmat_ptr(T)))
        return U, D, T

    def decompose_udt_prepivoted(self, A):
        """A = U diag(D) T by the stack's blocked QR with pre-pivoting (udt route 0)."""
        n = self.p.n
        A = _as_fcomplex(A, n)
        U = np.zeros((n, n), dtype=np.complex128, order="F")
        T = np.zeros((n, n), dtype=np.complex128, order="F")
        D = np.zeros(n)
        check(self.lib.dqmc_udt_prepivoted(self.h, _cmat_ptr(A), _cmat_ptr(U), dptr(D), _cmat_ptr(T)))
        return U, D, T

    def calculate_greens(self, Ul, Dl, Tl, Ur, Dr, Tr):
        """calculate_greens + calculate_logdet (stack.jl:338-385) on given factors."""
        n = self.p.n
        args = [_as_fcomplex(Ul, n), np.ascontiguousarray(Dl, dtype=np.float64), _as_fcomplex(Tl, n),
                _as_fcomplex(Ur, n), np.ascontiguousarray(Dr, dtype=np.float64), _as_fcomplex(Tr, n)]
        G = np.zeros((n, n), dtype=np.complex128, order="F")
        ld = ctypes.c_double()
        check(self.lib.dqmc_calculate_greens(self.h, _cmat_ptr(args[0]), dptr(args[1]), _cmat_ptr(args[2]),
                                             _cmat_ptr(args[3]), dptr(args[4]), _cmat_ptr(args[5]),
                                             _cmat_ptr(G), ctypes.byref(ld)))
        return G, ld.value

    def calc_greens_and_logdet(self, slice_):
        """fermion_measurements.jl:1068-1101 (G(slice) from scratch)."""
        n = self.p.n
        G = np.zeros((n, n), dtype=np.complex128, order="F")
        ld = ctypes.c_double()
        check(self.lib.dqmc_greens_at(self.h, slice_, _cmat_ptr(G), ctypes.byref(ld)))
        return G, ld.value

    def calc_detratio(self, site, new_op):
        """calc_detratio (local_updates.jl:42-59): (det, delta, M)."""
        op = np.ascontiguousarray(new_op, dtype=np.float64)
        det = np.zeros(2)
        delta = np.zeros((4, 4), dtype=np.complex128, order="F")
        M = np.zeros((4, 4), dtype=np.complex128, order="F")
        check(self.lib.dqmc_detratio(self.h, site, dptr(op), dptr(det), _cmat_ptr(delta), _cmat_ptr(M)))
        return complex(det[0], det[1]), delta, M

    def update_greens(self, site):
        """update_greens! (local_updates.jl:61-95) with the last calc_detratio's delta and M."""
        check(self.lib.dqmc_update_greens(self.h, site))

    def gemm(self, A, B, opA=0, opB=0):
        """mul!(C, op(A), op(B)) on the device; op 1 = adjoint."""
        n = self.p.n
        A = _as_fcomplex(A, n)
        B = _as_fcomplex(B, n)
        C = np.zeros((n, n), dtype=np.complex128, order="F")
        check(self.lib.dqmc_gemm(self.h, opA, opB, _cmat_ptr(A), _cmat_ptr(B), _cmat_ptr(C)))
        return C

    def inv(self, A):
        """(inv(A), log|det A|) on the device."""
        n = self.p.n
        A = _as_fcomplex(A, n)
        X = np.zeros((n, n), dtype=np.complex128, order="F")
        ld = ctypes.c_double()
        check(self.lib.dqmc_inverse(self.h, _cmat_ptr(A), _cmat_ptr(X), ctypes.byref(ld)))
        return X, ld.value

    def calc_boson_action(self):
        s = ctypes.c_double()
        check(self.lib.dqmc_boson_action(self.h, ctypes.byref(s)))
        return s.value

    # -- measurement hand-off (fermion_measurements.jl:1125-1193, 1239-1576) ---------------
    def effective_greens2greens(self, G):
        """effective_greens2greens (fermion_measurements.jl:1204-1208): a new matrix."""
        X = _as_fcomplex(G, self.p.n).copy(order="F")
        check(self.lib.dqmc_effective_greens2greens(self.h, _cmat_ptr(X)))
        return X

    def allocate_tdgfs(self):
        """allocate_tdgfs! (fermion_measurements.jl:1239-1265): device storage."""
        check(self.lib.dqmc_tdgf_alloc(self.h))

    def deallocate_tdgfs(self):
        check(self.lib.dqmc_tdgf_free(self.h))

    def measure_tdgfs(self):
        """measure_tdgfs! (fermion_measurements.jl:1343-1407); results stay on the device,
        read them with tdgf(tau)."""
        check(self.lib.dqmc_measure_tdgfs(self.h))

    def tdgf(self, tau):
        """(Gt0[tau], G0t[tau]) for 1 <= tau <= M."""
        n = self.p.n
        a = np.zeros((n, n), dtype=np.complex128, order="F")
        b = np.zeros((n, n), dtype=np.complex128, order="F")
        check(self.lib.dqmc_get_tdgf(self.h, tau, _cmat_ptr(a), _cmat_ptr(b)))
        return a, b

    TDGF_STACKS = ("BT0Inv", "BBetaT", "BT0", "BBetaTInv")

    def tdgf_stack(self, name, k):
        """Entry k (0-based) of one of the four checkpoint chains as (U, D, T)."""
        n = self.p.n
        U = np.zeros((n, n), dtype=np.complex128, order="F")
        T = np.zeros((n, n), dtype=np.complex128, order="F")
        D = np.zeros(n)
        check(self.lib.dqmc_get_tdgf_stack(self.h, self.TDGF_STACKS.index(name), k, _cmat_ptr(U), dptr(D),
                                           _cmat_ptr(T)))
        return U, D, T

    def calc_tdgf_direct(self, slice_):
        """calc_tdgf_direct (fermion_measurements.jl:1553-1576)."""
        n = self.p.n
        G = np.zeros((n, n), dtype=np.complex128, order="F")
        check(self.lib.dqmc_tdgf_direct(self.h, slice_, _cmat_ptr(G)))
        return G

    # -- fermionic observables (fermion_measurements.jl:216-229, 363-957) ------------------
    def measure_greens(self):
        """measure_greens (fermion_measurements.jl:225-229): G(1) from scratch, effective -> true;
        also kept on the device as the equal-time Green's function of the observables."""
        G = np.zeros((self.p.n, self.p.n), dtype=np.complex128, order="F")
        check(self.lib.dqmc_measure_greens(self.h, _cmat_ptr(G)))
        return G

    def measure_current_greens(self):
        """The greens observable of _measure_observables! (dqmc_framework.jl:378-382): the
        chain's G wrapped up from current_slice, effective -> true (kept on the device too)."""
        G = np.zeros((self.p.n, self.p.n), dtype=np.complex128, order="F")
        check(self.lib.dqmc_measure_current_greens(self.h, _cmat_ptr(G)))
        return G

    def set_etgf(self, G):
        """Use a host equal-time Green's function for the observables."""
        G = _as_fcomplex(G, self.p.n)
        check(self.lib.dqmc_set_etgf(self.h, _cmat_ptr(G)))

    def _obs_pair(self, fn, real):
        L = self.p.L
        if real:
            a, b = np.zeros((L, L), order="F"), np.zeros((L, L), order="F")
            check(fn(self.h, dptr(a), dptr(b)))
        else:
            a = np.zeros((L, L), dtype=np.complex128, order="F")
            b = np.zeros((L, L), dtype=np.complex128, order="F")
            check(fn(self.h, _cmat_ptr(a), _cmat_ptr(b)))
        return a, b

    def measure_etpc(self):
        """measure_etpc! (fermion_measurements.jl:388-470): (etpc_plus, etpc_minus), L x L [y, x]."""
        return self._obs_pair(self.lib.dqmc_measure_etpc, True)

    def measure_zfpc(self):
        """measure_zfpc! (fermion_measurements.jl:500-562) on the last measure_tdgfs."""
        return self._obs_pair(self.lib.dqmc_measure_zfpc, True)

    def measure_etcdc(self):
        """measure_etcdc! (fermion_measurements.jl:592-657): (etcdc_plus, etcdc_minus)."""
        return self._obs_pair(self.lib.dqmc_measure_etcdc, False)

    def measure_zfcdc(self):
        """measure_zfcdc! (fermion_measurements.jl:688-778) with greens = the etgf."""
        return self._obs_pair(self.lib.dqmc_measure_zfcdc, False)

    def measure_zfccc(self):
        """measure_zfccc! (fermion_measurements.jl:814-899) with greens = the etgf, and
        measure_sfdensity (:911-923): (zfccc_xx, zfccc_yy, sfdensity)."""
        if not getattr(self, "_hopping_set", False):
            from .model import hopping_matrix
            T = hopping_matrix(self.p)
            check(self.lib.dqmc_set_hopping_matrix(self.h, _cmat_ptr(T)))
            self._hopping_set = True
        L = self.p.L
        xx = np.zeros((L, L), dtype=np.complex128, order="F")
        yy = np.zeros((L, L), dtype=np.complex128, order="F")
        rho = ctypes.c_double()
        check(self.lib.dqmc_measure_zfccc(self.h, _cmat_ptr(xx), _cmat_ptr(yy), ctypes.byref(rho)))
        return xx, yy, rho.value

    def grad(self):
        """dS/dphi (3, N, M) and S = S_B + log|det G| of the current field."""
        g = np.zeros((3, self.p.N, self.p.slices), order="F")
        s = ctypes.c_double()
        check(self.lib.dqmc_grad(self.h, dptr(g), ctypes.byref(s)))
        return g, s.value


class Batch:
    """Many independent chains (possibly of different L, M) advanced together."""

    def __init__(self, params_list, device=0, streams=None):
        self.lib = _capi.load()
        self.params = list(params_list)
        self._keep = []
        nc = len(self.params)
        P = (_capi.Params * nc)(*[_params_struct(p) for p in self.params])
        O = (_capi.Operators * nc)(*[_operators_struct(build_operators(p), self._keep) for p in self.params])
        h = ctypes.c_void_p()
        check(self.lib.dqmc_batch_create(P, O, nc, device, ctypes.byref(h)))
        self.h = h
        self.streams = streams if streams is not None else [UniformStream(p.seed) for p in self.params]
        self._pending_uniforms = False

    def close(self):
        if getattr(self, "h", None):
            self.lib.dqmc_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        if sys.is_finalizing():
            return
        try:
            self.close()
        except Exception:
            pass

    def set_field(self, chain, phi):
        phi = np.asfortranarray(phi, dtype=np.float64)
        check(self.lib.dqmc_batch_set_field(self.h, chain, dptr(phi)))

    def get_field(self, chain):
        p = self.params[chain]
        phi = np.zeros((3, p.N, p.slices), order="F")
        check(self.lib.dqmc_batch_get_field(self.h, chain, dptr(phi)))
        return phi

    def init(self, confs):
        for c, conf in enumerate(confs):
            self.set_field(c, conf)
        check(self.lib.dqmc_batch_init(self.h))

    def sweep(self, nsweeps=1, sync=True):
        self.synchronize()  # consumes the uniforms of an earlier unsynchronised sweep
        for c, p in enumerate(self.params):
            vals = np.ascontiguousarray(self.streams[c].take(4 * p.N * p.slices * nsweeps))
            check(self.lib.dqmc_batch_set_uniforms(self.h, c, dptr(vals), len(vals)))
        box = np.array([p.box for p in self.params])
        check(self.lib.dqmc_batch_sweep(self.h, nsweeps, dptr(box)))
        self._pending_uniforms = True
        if sync:
            self.synchronize()

    def synchronize(self):
        """Waits for every chain's queued work; the uniforms an unsynchronised sweep drew are
        then consumed from the host streams."""
        check(self.lib.dqmc_batch_synchronize(self.h))
        if self._pending_uniforms:
            for c in range(len(self.params)):
                self.streams[c].advance(self.stats(c).uniforms_used)
            self._pending_uniforms = False

    def set_groups(self, wg_qr=0, wg_inv=0, wg_local=0):
        """Workgroups per chain of every shape group (0 keeps the current value)."""
        check(self.lib.dqmc_batch_set_groups(self.h, wg_qr, wg_inv, wg_local))

    def set_local_variant(self, variant):
        """Local-update variant of every shape group (see DQMC.set_local_variant)."""
        check(self.lib.dqmc_batch_set_local_variant(self.h, variant))

    def use_graph(self, enable=True):
        """Capture each shape group's sweep in a HIP graph and replay it (see DQMC.use_graph)."""
        check(self.lib.dqmc_batch_use_graph(self.h, 1 if enable else 0))

    def set_merged(self, enable=True):
        """Sweep all shape groups in lockstep, one launch per kernel for every chain (default),
        or group by group."""
        check(self.lib.dqmc_batch_set_merged(self.h, 1 if enable else 0))

    def autotune(self, reps=3):
        """DQMC.autotune for every shape group; returns one dict per group."""
        box = np.array([p.box for p in self.params])
        rep = np.zeros(6 * len(self.params))
        check(self.lib.dqmc_batch_autotune(self.h, dptr(box), reps, dptr(rep)))
        # shape groups, in order of first appearance (dqmc_batch_create)
        keys = []
        for p in self.params:
            k = (p.L, p.slices, p.safe_mult, bool(p.checkerboard), bool(p.all_checks))
            if k not in keys:
                keys.append(k)
        return [_tune_report(rep[6 * g:6 * g + 6]) for g in range(len(keys))]

    def greens(self, chain):
        n = self.params[chain].n
        G = np.zeros((n, n), dtype=np.complex128, order="F")
        check(self.lib.dqmc_batch_get_greens(self.h, chain, _cmat_ptr(G)))
        return G

    def stats(self, chain):
        st = _capi.Stats()
        check(self.lib.dqmc_batch_get_stats(self.h, chain, ctypes.byref(st)))
        return st


def _tune_report(r):
    return {"local_workgroups": int(r[0]), "local_variant": int(r[1]), "local_ms": float(r[2]),
            "qr_workgroups": int(r[3]), "qr_panel": int(r[4]), "qr_ms": float(r[5])}


def _hmc_trajectory(self, momenta, nsteps, eps, u):
    """One HMC trajectory on the device; returns (accepted, dH)."""
    mom = np.asfortranarray(momenta, dtype=np.float64)
    if mom.shape != (3, self.p.N, self.p.slices):
        raise ValueError("momenta must have shape (3, N, M)")
    acc = ctypes.c_int()
    dh = ctypes.c_double()
    check(self.lib.dqmc_hmc_trajectory(self.h, dptr(mom), nsteps, eps, u, ctypes.byref(acc), ctypes.byref(dh)))
    return bool(acc.value), dh.value


DQMC.hmc_trajectory = _hmc_trajectory
