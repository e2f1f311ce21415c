// Local updates of one time slice: N sequential Metropolis proposals, each a
// 4x4 determinant ratio, and on acceptance a rank-4 update of G.
//
// Reference: local_updates (local_updates.jl:1-39), calc_detratio (:42-59),
// update_greens! (:61-95), calc_boson_action_diff (action.jl:57-101),
// randuniform (dqmc_framework.jl:628-635).
//
// One chain group of W workgroups per chain (cooperative launch). Every
// workgroup evaluates every proposal redundantly from identical inputs (the
// host uniform stream, an LDS copy of the field slice, the 4x4 block of G), so
// all reach the same decision without communicating; rejected proposals cost no
// synchronisation. Workgroup w owns a contiguous block of G's columns. On
// acceptance each workgroup reads the four site columns of G (the rank-4 left
// factor) and the four site rows of its own columns, and writes its columns of
// the updated G into the ping-pong partner buffer, so no workgroup ever writes
// memory another is reading in the same step: one group barrier per accepted
// proposal.
#include <vector>

#include "chain.hpp"

namespace dqmc {

// coefficients of interaction_matrix_exp_op (interactions.jl:102-140): C = cosh(x),
// sh = power sinh(x) / |o|, x = lam dtau |o| >= 0. Both from one expm1 and one reciprocal
// (em = e^x - 1: cosh x = 1 + em^2 / (2 e^x), sinh x = em (1 + e^-x) / 2), accurate for
// small x too; this sits on the critical path of every proposal.
__device__ inline void ev_coeffs_op(double lam_dtau, double o0, double o1, double o2, double power, double& C,
                                    cplx& S, double& R) {
  const double nrm = sqrt(o0 * o0 + o1 * o1 + o2 * o2);
  const double em = expm1(lam_dtau * nrm);
  const double ie = 1.0 / (1.0 + em);
  C = fma(0.5 * em, em * ie, 1.0);
  const double sh = power * (0.5 * em * (1.0 + ie)) / nrm;
  S = mk(-o0 * sh, o1 * sh);
  R = -o2 * sh;
}

__device__ inline void ev_op(double lam_dtau, double o0, double o1, double o2, double power, cplx E[16]) {
  double C, R;
  cplx S;
  ev_coeffs_op(lam_dtau, o0, o1, o2, power, C, S, R);
  const cplx z = mk(0, 0);
  E[0] = mk(C, 0);   E[1] = S;          E[2] = z;          E[3] = mk(R, 0);
  E[4] = conj(S);    E[5] = mk(C, 0);   E[6] = mk(-R, 0);  E[7] = z;
  E[8] = z;          E[9] = mk(-R, 0);  E[10] = mk(C, 0);  E[11] = conj(S);
  E[12] = mk(R, 0);  E[13] = z;         E[14] = S;         E[15] = mk(C, 0);
}

// row-major 4x4 product
__device__ inline void mul4(const cplx* A, const cplx* B, cplx* C) {
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      cplx acc = mk(0, 0);
#pragma unroll
      for (int k = 0; k < 4; ++k) cfma(acc, A[i * 4 + k], B[k * 4 + j]);
      C[i * 4 + j] = acc;
    }
}

__device__ inline double cabs1_(cplx a) { return fabs(a.re) + fabs(a.im); }

// det(M) by LU with partial pivoting (first max of |re|+|im|), as LinearAlgebra.det
__device__ inline cplx det4(const cplx* Min) {
  cplx a[16];
#pragma unroll
  for (int e = 0; e < 16; ++e) a[e] = Min[e];
  cplx det = mk(1, 0);
  for (int k = 0; k < 4; ++k) {
    int p = k;
    double bv = cabs1_(a[k * 4 + k]);
    for (int r = k + 1; r < 4; ++r) {
      const double v = cabs1_(a[r * 4 + k]);
      if (v > bv) {
        bv = v;
        p = r;
      }
    }
    if (p != k) {
      for (int j = 0; j < 4; ++j) {
        const cplx t = a[k * 4 + j];
        a[k * 4 + j] = a[p * 4 + j];
        a[p * 4 + j] = t;
      }
      det = -det;
    }
    const cplx piv = a[k * 4 + k];
    det = det * piv;
    const cplx pinv = cinv(piv);
    for (int r = k + 1; r < 4; ++r) {
      const cplx l = a[r * 4 + k] * pinv;
      for (int j = k + 1; j < 4; ++j) a[r * 4 + j] = a[r * 4 + j] - l * a[k * 4 + j];
    }
  }
  return det;
}

// inverse by Gauss-Jordan with partial pivoting
__device__ inline void inv4(const cplx* Min, cplx* out) {
  cplx a[16];
  int piv[4];
#pragma unroll
  for (int e = 0; e < 16; ++e) a[e] = Min[e];
  for (int k = 0; k < 4; ++k) {
    int p = k;
    double bv = cabs1_(a[k * 4 + k]);
    for (int r = k + 1; r < 4; ++r) {
      const double v = cabs1_(a[r * 4 + k]);
      if (v > bv) {
        bv = v;
        p = r;
      }
    }
    piv[k] = p;
    if (p != k)
      for (int j = 0; j < 4; ++j) {
        const cplx t = a[k * 4 + j];
        a[k * 4 + j] = a[p * 4 + j];
        a[p * 4 + j] = t;
      }
    const cplx pinv = cinv(a[k * 4 + k]);
    a[k * 4 + k] = mk(1, 0);
    for (int j = 0; j < 4; ++j) a[k * 4 + j] = a[k * 4 + j] * pinv;
    for (int r = 0; r < 4; ++r) {
      if (r == k) continue;
      const cplx f = a[r * 4 + k];
      a[r * 4 + k] = mk(0, 0);
      for (int j = 0; j < 4; ++j) a[r * 4 + j] = a[r * 4 + j] - f * a[k * 4 + j];
    }
  }
  for (int k = 3; k >= 0; --k)
    if (piv[k] != k)
      for (int r = 0; r < 4; ++r) {
        const cplx t = a[r * 4 + k];
        a[r * 4 + k] = a[r * 4 + piv[k]];
        a[r * 4 + piv[k]] = t;
      }
#pragma unroll
  for (int e = 0; e < 16; ++e) out[e] = a[e];
}

struct LocalShared {
  cplx delta[16];
  cplx minv[16];
  int site;  // site of the accepted proposal to apply, N when the slice is done
  int stop;  // the uniform stream ran out
};

// Site block [[C,S,0,R],[S*,C,-R,0],[0,-R,C,S*],[R,0,S,C]] (interactions.jl:102-140): type of the
// entry at row-major position e: 0 zero, 1 C, 2 S, 3 conj(S), 4 R, 5 -R.
__device__ inline int ev_type(int e) {
  const unsigned long long code = 0x1204315005134021ull;  // nibble e = type of entry e (e = 0 lowest)
  return (int)((code >> (4 * e)) & 0xF);
}
__device__ inline cplx ev_entry_t(double C, cplx S, double R, int t) {
  const double re = t == 1 ? C : (t == 2 || t == 3) ? S.re : t == 4 ? R : t == 5 ? -R : 0.0;
  const double im = t == 2 ? S.im : t == 3 ? -S.im : 0.0;
  return mk(re, im);
}

// determinant of a 4x4 by Laplace expansion along rows (0, 1) and (2, 3)
__device__ inline cplx det4_laplace(const cplx* m) {
  auto m2 = [&](int r, int p, int q) { return m[r * 4 + p] * m[(r + 1) * 4 + q] - m[r * 4 + q] * m[(r + 1) * 4 + p]; };
  cplx d = m2(0, 0, 1) * m2(2, 2, 3);
  d = d - m2(0, 0, 2) * m2(2, 1, 3);
  d = d + m2(0, 0, 3) * m2(2, 1, 2);
  d = d + m2(0, 1, 2) * m2(2, 0, 3);
  d = d - m2(0, 1, 3) * m2(2, 0, 2);
  d = d + m2(0, 2, 3) * m2(2, 0, 1);
  return d;
}

__device__ inline cplx shfl_c(cplx v, int src) {
  return mk(__shfl(v.re, src, 64), __shfl(v.im, src, 64));
}
// DPP move of a complex value (quad_perm 0x00-0xFF, row_ror 0x121-0x12F, row_newbcast 0x150-0x15F)
template <int CTRL>
__device__ inline cplx dpp_c(cplx v) {
  return mk(dpp_move<CTRL>(v.re), dpp_move<CTRL>(v.im));
}

// ---------------------------------------------------------------------------
// Shared pieces of both local-update kernels
// ---------------------------------------------------------------------------

// LDS view of one slice: field of this and the two neighbouring slices, neighbour
// table, the slice's uniforms, the old-field site blocks eV_op(phi_i, -1), and per site
// the sums of the field over its two time neighbours and its four space neighbours
// (the latter kept current as sites are accepted).
struct SliceLds {
  double* phl;  // [3N] current slice
  double* phe;  // [3N] earlier slice
  double* pla;  // [3N] later slice
  int* nbl;     // [4N]
  double* ul;   // [4N]
  double* e1c;  // [4N] C, S.re, S.im, R
  double* tnb;  // [3N] pla + phe
  double* snb;  // [3N] sum over the four space neighbours of phl
  long long cur0, navail;
};
// doubles of LDS SliceLds takes (rounded up to keep 16-byte alignment after it)
__host__ __device__ inline size_t slice_lds_doubles(int N) { return (size_t)25 * N + (N & 1); }

__device__ inline SliceLds load_slice(const ChainDesc& c, int l, double* smem) {
  const int N = c.N, M = c.M, tid = threadIdx.x;
  SliceLds s;
  s.phl = smem;
  s.phe = s.phl + 3 * N;
  s.pla = s.phe + 3 * N;
  s.nbl = reinterpret_cast<int*>(s.pla + 3 * N);
  s.ul = smem + 11 * N;
  s.e1c = s.ul + 4 * N;
  s.tnb = s.e1c + 4 * N;
  s.snb = s.tnb + 3 * N;
  const int le = (l == 0) ? M - 1 : l - 1;  // time_neighbors (lattice.jl:144-153)
  const int la = (l == M - 1) ? 0 : l + 1;
  for (int e = tid; e < 3 * N; e += blockDim.x) {
    s.phl[e] = c.phi[(size_t)3 * N * l + e];
    s.phe[e] = c.phi[(size_t)3 * N * le + e];
    s.pla[e] = c.phi[(size_t)3 * N * la + e];
  }
  for (int e = tid; e < 4 * N; e += blockDim.x) s.nbl[e] = c.nb[e];
  // neighbour sums in the order calc_boson_action_diff uses (action.jl:57-101), from
  // global memory so no barrier is needed before them
  for (int e = tid; e < 3 * N; e += blockDim.x) {
    const int i = e / 3, a = e - 3 * i;
    s.tnb[e] = c.phi[(size_t)3 * N * la + e] + c.phi[(size_t)3 * N * le + e];
    double sn = 0.0;
    for (int k = 0; k < 4; ++k) sn += c.phi[(size_t)3 * N * l + 3 * c.nb[k * N + i] + a];
    s.snb[e] = sn;
  }
  // at most 4 draws per proposal: this slice consumes no more than 4N uniforms
  s.cur0 = *c.ucursor;
  s.navail = min((long long)4 * N, c.n_uniforms - s.cur0);
  for (long long e = tid; e < s.navail; e += blockDim.x) s.ul[e] = c.uniforms[s.cur0 + e];
  // each site is visited once per slice, so its old field is the value at kernel start
  const double lam_dtau = c.lam * c.dtau;
  for (int i = tid; i < N; i += blockDim.x) {
    const double* f = c.phi + (size_t)3 * N * l + 3 * i;
    double C, R;
    cplx S;
    ev_coeffs_op(lam_dtau, f[0], f[1], f[2], -1.0, C, S, R);
    s.e1c[4 * i + 0] = C;
    s.e1c[4 * i + 1] = S.re;
    s.e1c[4 * i + 2] = S.im;
    s.e1c[4 * i + 3] = R;
  }
  return s;
}

// Wave-0 state of the proposal sequence of one slice.
struct Proposals {
  long long cur, logi;
  double dS_acc, minratio;
  int nacc, nonreal, i;
  // an accepted proposal whose field update is deferred (commit_accept): site (-1: none),
  // exp(-dS) and the new field
  int pend_site;
  double pend_e, pend_n0, pend_n1, pend_n2;
};

// The accepted proposal's field update (wave 0): the field of the site, the four neighbours'
// space sums recomputed in the original order (action.jl:57-101), and the boson action
// p.boson_action += -log(exp_delta_S_boson) (local_updates.jl:34). The next evaluation reads
// the field and sums, nothing before it does.
__device__ inline void commit_accept(const ChainDesc& c, const SliceLds& s, Proposals& p) {
  const int site = p.pend_site;
  if (site < 0) return;
  const int N = c.N, lane = threadIdx.x & 63;
  if (lane == 0) {
    s.phl[3 * site + 0] = p.pend_n0;
    s.phl[3 * site + 1] = p.pend_n1;
    s.phl[3 * site + 2] = p.pend_n2;
  }
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");  // lane 0's field before the sums read it
  if (lane < 12) {
    const int j = s.nbl[(lane / 3) * N + site], a = lane - 3 * (lane / 3);
    double sn = 0.0;
    for (int k = 0; k < 4; ++k) sn += s.phl[3 * s.nbl[k * N + j] + a];
    s.snb[3 * j + a] = sn;
  }
  p.dS_acc += -log(p.pend_e);
  p.pend_site = -1;
}

// Evaluate proposals from site p.i on, in wave 0 (each 16-lane row of the wave holds one
// site's 4x4 algebra, lane (ea, eb) of the row one entry), until one is accepted: its delta
// and M^-1 go to sh.delta / sh.minv, the field of the site is updated in LDS and the site is
// returned.
// Sites from p.i up to limit - 1 are visited; returns limit when none of them is accepted
// (p.i = limit) or when the uniform stream ran out (*stop set, p.i = the site it stopped at).
// diag(j) returns this lane's entry G_jj(ea, eb) of the current G.
// Reference: local_updates (local_updates.jl:1-39), calc_detratio (:42-59),
// calc_boson_action_diff (action.jl:57-101).
template <class Diag>
__device__ int evaluate_proposals(const ChainDesc& c, const SliceLds& s, Proposals& p, LocalShared& sh,
                                  double box, int dolog, bool logger, Diag diag, int* stop, int limit,
                                  bool defer = false) {
  const int N = c.N, lane = threadIdx.x & 63;
  const int row = lane >> 4;  // 16-lane row of the wave: it evaluates site i + row
  const int ea = (lane & 15) >> 2, eb = lane & 3;
  int t1[4], t2[4];  // entry types of row ea (old operator) and column (ea + eb) & 3 (new operator)
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    t1[k] = ev_type(ea * 4 + k);
    t2[k] = ev_type(k * 4 + ((ea + eb) & 3));
  }
  const double lam_dtau = c.lam * c.dtau;
  const double k_time = 1.0 / (c.dtau * c.c * c.c), k_space = 0.5 * c.dtau, k_dt = c.dtau;
  const double k_r = 0.5 * c.r, k_u = 0.25 * c.u;
  const int edrun = c.edrun;
  *stop = 0;
  // The four rows evaluate four consecutive sites at once. Row r takes site i + r with the
  // uniforms it gets if sites i .. i+r-1 are all rejected -- a rejection needs p_acc <= 1 and
  // so always draws 4 -- from the current G and field, which those rejections leave as they
  // are. The rows are then resolved in order, exactly as one proposal after another; the
  // first acceptance ends the call and the later rows' work is dropped.
  int i = p.i;
  cplx gnext = mk(0, 0);  // this row's G_ii, loaded one step ahead
  if (i < limit) gnext = diag(min(i + row, limit - 1));
  for (; i < limit; i += 4) {
    const int si = min(i + row, limit - 1);
    const cplx gv = gnext;
    if (i + 4 < limit) gnext = diag(min(i + 4 + row, limit - 1));
    const long long off = min((p.cur - s.cur0) + 4 * row, max(s.navail - 4, 0ll));  // in bounds
    const double* u = s.ul + off;
    double old[3], nw[3];
    for (int a = 0; a < 3; ++a) {
      old[a] = s.phl[3 * si + a];
      nw[a] = old[a] + (-box + 2.0 * box * u[a]);
    }
    const double old_sq = old[0] * old[0] + old[1] * old[1] + old[2] * old[2];
    const double new_sq = nw[0] * nw[0] + nw[1] * nw[1] + nw[2] * nw[2];
    const double sq_diff = new_sq - old_sq;
    const double pow4_diff = new_sq * new_sq - old_sq * old_sq;
    double dt_dot = 0.0, ds_dot = 0.0;
    for (int a = 0; a < 3; ++a) {
      const double diff = nw[a] - old[a];
      dt_dot += s.tnb[3 * si + a] * diff;
      ds_dot += s.snb[3 * si + a] * diff;
    }
    double dS;
    if (!edrun) {
      dS = k_time * (sq_diff - dt_dot);
      dS += k_space * (4 * sq_diff - 2.0 * ds_dot);
      dS += k_dt * (k_r * sq_diff + k_u * pow4_diff);
    } else {
      dS = k_dt * (k_r * sq_diff);
    }
    const double exp_dS = exp(-dS);
    // delta = eV_op(old, -1) eV_op(new, +1) - I,  M = I + delta (I - G_ii)
    double C2, R2;
    cplx S2;
    ev_coeffs_op(lam_dtau, nw[0], nw[1], nw[2], 1.0, C2, S2, R2);
    const double C1 = s.e1c[4 * si], R1 = s.e1c[4 * si + 3];
    const cplx S1 = mk(s.e1c[4 * si + 1], s.e1c[4 * si + 2]);
    // this lane's entry of delta is (ea, (ea + eb) & 3): the skew lets the product below
    // take both operands from DPP moves (quad broadcasts of delta, row rotations of I - G);
    // every DPP move stays inside the 16-lane row
    cplx dm = mk(0, 0);
#pragma unroll
    for (int k = 0; k < 4; ++k) cfma(dm, ev_entry_t(C1, S1, R1, t1[k]), ev_entry_t(C2, S2, R2, t2[k]));
    if (eb == 0) dm.re -= 1.0;
    const cplx gm = mk((ea == eb ? 1.0 : 0.0) - gv.re, -gv.im);
    // mm(ea, eb) = sum_j delta(ea, (ea - j) & 3) (I - G)((ea - j) & 3, eb)
    cplx mm = mk(0, 0);
    cfma(mm, dpp_c<0x00>(dm), gm);
    cfma(mm, dpp_c<0xFF>(dm), dpp_c<0x124>(gm));
    cfma(mm, dpp_c<0xAA>(dm), dpp_c<0x128>(gm));
    cfma(mm, dpp_c<0x55>(dm), dpp_c<0x12C>(gm));
    if (ea == eb) mm.re += 1.0;
    cplx Mall[16];  // every lane: all of its row's M, by row broadcasts of each lane of the row
    Mall[0] = dpp_c<0x150>(mm);
    Mall[1] = dpp_c<0x151>(mm);
    Mall[2] = dpp_c<0x152>(mm);
    Mall[3] = dpp_c<0x153>(mm);
    Mall[4] = dpp_c<0x154>(mm);
    Mall[5] = dpp_c<0x155>(mm);
    Mall[6] = dpp_c<0x156>(mm);
    Mall[7] = dpp_c<0x157>(mm);
    Mall[8] = dpp_c<0x158>(mm);
    Mall[9] = dpp_c<0x159>(mm);
    Mall[10] = dpp_c<0x15A>(mm);
    Mall[11] = dpp_c<0x15B>(mm);
    Mall[12] = dpp_c<0x15C>(mm);
    Mall[13] = dpp_c<0x15D>(mm);
    Mall[14] = dpp_c<0x15E>(mm);
    Mall[15] = dpp_c<0x15F>(mm);
    const cplx det = det4_laplace(Mall);
    const double p_row = exp_dS * det.re;
    const double u3 = u[3];
    // resolve the rows in order (wave-uniform)
    for (int g = 0; g < 4; ++g) {
      const int site = i + g;
      if (site >= limit) break;
      if ((p.cur - s.cur0) + 4 > s.navail) {
        *stop = 1;
        p.i = site;
        return limit;
      }
      const double p_acc = read_lane(p_row, 16 * g);
      const double dre = read_lane(det.re, 16 * g), dim = read_lane(det.im, 16 * g);
      if (fabs(dim) > 1e-4 * fabs(dre)) ++p.nonreal;
      int acc;
      double uval = -1.0;
      if (p_acc > 1.0) {
        acc = 1;
        p.cur += 3;
      } else {
        uval = read_lane(u3, 16 * g);
        acc = uval < p_acc;
        p.cur += 4;
        p.minratio = fmin(p.minratio, fabs(p_acc - uval));
      }
      if (dolog && logger && lane == 0 && p.logi < c.log_cap) {
        double* lg = c.prop_log + 4 * p.logi;
        lg[0] = p_acc;
        lg[1] = uval;
        lg[2] = acc;
        lg[3] = dim / dre;
      }
      ++p.logi;
      if (acc) {
        ++p.nacc;
        p.pend_site = site;
        p.pend_e = read_lane(exp_dS, 16 * g);
        p.pend_n0 = read_lane(nw[0], 16 * g);
        p.pend_n1 = read_lane(nw[1], 16 * g);
        p.pend_n2 = read_lane(nw[2], 16 * g);
        if (!defer) commit_accept(c, s, p);
        // M^-1 (b, a) = cofactor (a, b) / det M, one entry per lane of row g
        cplx m3[9];
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
          for (int q = 0; q < 3; ++q) m3[r * 3 + q] = shfl_c(mm, 16 * g + (r + (r >= ea)) * 4 + q + (q >= eb));
        cplx cof = m3[0] * (m3[4] * m3[8] - m3[5] * m3[7]) - m3[1] * (m3[3] * m3[8] - m3[5] * m3[6]) +
                   m3[2] * (m3[3] * m3[7] - m3[4] * m3[6]);
        if ((ea + eb) & 1) cof = -cof;
        if (row == g) {
          sh.delta[ea * 4 + ((ea + eb) & 3)] = dm;
          sh.minv[eb * 4 + ea] = cof * cinv(det);
        }
        p.i = site + 1;
        return site;
      }
    }
  }
  p.i = limit;
  return limit;
}

// counters and the updated field slice, written by workgroup 0
__device__ inline void finish_slice(const ChainDesc& c, int l, const SliceLds& s, const Proposals& p, int dolog,
                                    int stop) {
  const int N = c.N, tid = threadIdx.x;
  __syncthreads();
  for (int e = tid; e < 3 * N; e += blockDim.x) c.phi[(size_t)3 * N * l + e] = s.phl[e];
  if (tid == 0) {
    *c.ucursor = p.cur;
    if (dolog) *c.log_count = p.logi;
    c.scal[S_BOSON_ACTION] += p.dS_acc;
    c.scal[S_ACC_SLICE] = p.nacc;
    c.scal[S_ACC_TOTAL] += p.nacc;
    c.scal[S_PROP_TOTAL] += N;
    c.scal[S_NONREAL] += p.nonreal;
    c.scal[S_MINRATIO] = p.minratio;
    if (stop) *c.error = 2;
  }
}

// ---------------------------------------------------------------------------
// Variant 1: G in global memory, ping-pong buffers (any size).
// Every workgroup evaluates every proposal redundantly; on acceptance each
// workgroup reads the site's four columns, updates its own column block into the
// partner buffer; one group barrier per accepted proposal.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_local_updates(const ChainDesc* chains, int slice, const int* slices,
                                                       double box, int dolog) {
  const ChainDesc& c = chains[blockIdx.y];
  const int n = c.n, N = c.N;
  const int W = gridDim.x, w = blockIdx.x, tid = threadIdx.x;
  const int l = slices ? slices[blockIdx.y] : slice;
  GroupSync g = chain_group(c);
  const int cb0 = (int)(((long long)n * w) / W), cb1 = (int)(((long long)n * (w + 1)) / W);
  const int ncol = cb1 - cb0;

  extern __shared__ double smem[];
  const SliceLds s = load_slice(c, l, smem);
  cplx* Acol = reinterpret_cast<cplx*>(smem + slice_lds_doubles(N));  // [n][4]
  cplx* Brow = Acol + 4 * (size_t)n;                                  // [4][ncol]
  __shared__ LocalShared sh;
  __syncthreads();

  Proposals p{s.cur0, dolog ? *c.log_count : 0, 0.0, c.scal[S_MINRATIO], 0, 0, 0, -1, 0.0, 0.0, 0.0, 0.0};
  int buf = 0;
  cplx* Gb[2] = {c.mat[M_G], c.mat[M_GALT]};
  for (;;) {
    if (tid < 64) {
      int stop;
      const cplx* Gcur = Gb[buf];
      const int lane = tid & 63, ea = (lane & 15) >> 2, eb = lane & 3;
      const int site = evaluate_proposals(
          c, s, p, sh, box, dolog, w == 0, [&](int j) { return Gcur[(size_t)(j + eb * N) * n + (j + ea * N)]; },
          &stop, N);
      if (tid == 0) {
        sh.site = site;
        sh.stop = stop;
      }
    }
    __syncthreads();
    const int site = sh.site;
    if (site >= N) break;
    const cplx* Gc = Gb[buf];
    cplx* Gn = Gb[buf ^ 1];
    // update_greens! (local_updates.jl:61-95): A = (G[:, c_i] - E) M^-1 (n x 4)
    for (int r = tid; r < n; r += blockDim.x) {
      cplx g4[4];
      for (int k = 0; k < 4; ++k) {
        g4[k] = Gc[(size_t)(site + k * N) * n + r];
        if (r == site + k * N) g4[k].re -= 1.0;
      }
      for (int k = 0; k < 4; ++k) {
        cplx acc = mk(0, 0);
        for (int m = 0; m < 4; ++m) cfma(acc, g4[m], sh.minv[m * 4 + k]);
        Acol[4 * (size_t)r + k] = acc;
      }
    }
    // B = delta G[r_i, own columns]   (4 x ncol)
    for (int jl = tid; jl < ncol; jl += blockDim.x) {
      const int j = cb0 + jl;
      cplx g4[4];
      for (int k = 0; k < 4; ++k) g4[k] = Gc[(size_t)j * n + (site + k * N)];
      for (int k = 0; k < 4; ++k) {
        cplx acc = mk(0, 0);
        for (int m = 0; m < 4; ++m) cfma(acc, sh.delta[k * 4 + m], g4[m]);
        Brow[(size_t)k * ncol + jl] = acc;
      }
    }
    __syncthreads();
    for (size_t e = tid; e < (size_t)ncol * n; e += blockDim.x) {
      const int jl = (int)(e / n), r = (int)(e - (size_t)jl * n);
      const size_t o = (size_t)(cb0 + jl) * n + r;
      cplx acc = Gc[o];
      const cplx* a = Acol + 4 * (size_t)r;
      for (int k = 0; k < 4; ++k) cfma(acc, a[k], Brow[(size_t)k * ncol + jl]);
      Gn[o] = acc;
    }
    buf ^= 1;
    group_sync(g);
    if (group_failed(g)) return;
  }
  if (buf == 1) {
    for (size_t e = tid; e < (size_t)ncol * n; e += blockDim.x) {
      const size_t o = (size_t)cb0 * n + e;
      Gb[0][o] = Gb[1][o];
    }
  }
  if (w == 0) finish_slice(c, l, s, p, dolog, sh.stop);
}

// ---------------------------------------------------------------------------
// Variant 2: G resident in LDS. Workgroup w holds its column block of G in LDS for
// the whole slice and applies every accepted rank-4 update there. What the other
// workgroups need crosses global memory in "rounds" (publish, then one group
// barrier), double-buffered by round parity:
//   * the 4x4 diagonal blocks G_jj of every site not yet visited (decisions), and
//   * the full columns of the next kLocalWindow sites (the rank-4 left factor of
//     the next accepted proposal; a proposal accepted beyond the window costs one
//     extra round that publishes a window starting at it).
// Before the first round of a slice both come from G in global memory, which the
// tiles were loaded from. Global G is written back once, at the end.
// ---------------------------------------------------------------------------
constexpr int kLocalWindow = 4;

__global__ __launch_bounds__(256) void k_local_updates_tile(const ChainDesc* chains, int slice, const int* slices,
                                                            double box, int dolog) {
  const ChainDesc& c = chains[blockIdx.y];
  const int n = c.n, N = c.N;
  const int W = gridDim.x, w = blockIdx.x, tid = threadIdx.x;
  const int l = slices ? slices[blockIdx.y] : slice;
  GroupSync g = chain_group(c);
  const int cb0 = (int)(((long long)n * w) / W), cb1 = (int)(((long long)n * (w + 1)) / W);
  const int ncol = cb1 - cb0;

  extern __shared__ double smem[];
  const SliceLds s = load_slice(c, l, smem);
  cplx* T = reinterpret_cast<cplx*>(smem + slice_lds_doubles(N));  // [ncol][n] own columns of G
  cplx* Brow = T + (size_t)ncol * n;                                // [4][ncol]
  __shared__ LocalShared sh;
  __shared__ int pub_site;  // first site of the window published in the last round
  const cplx* G0 = c.mat[M_G];
  for (size_t e = tid; e < (size_t)ncol * n; e += blockDim.x) T[e] = G0[(size_t)cb0 * n + e];
  if (tid == 0) pub_site = -1;
  __syncthreads();

  const size_t win_sz = (size_t)kLocalWindow * 4 * n;  // complex entries of one window
  Proposals p{s.cur0, dolog ? *c.log_count : 0, 0.0, c.scal[S_MINRATIO], 0, 0, 0, -1, 0.0, 0.0, 0.0, 0.0};
  int rnd = 0;  // completed rounds; buffers of round r live at parity r & 1
  int stop = 0;

  // one round: publish the diagonal blocks of sites >= from and the columns of sites
  // win0 .. win0 + kLocalWindow - 1 from the own tile, then a group barrier
  auto publish_round = [&](int from, int win0) -> bool {
    const int par = (rnd + 1) & 1;
    cplx* dg = c.lu_diag + (size_t)par * 16 * N;
    cplx* wn = c.lu_win + (size_t)par * win_sz;
    for (int jl = tid; jl < ncol; jl += blockDim.x) {
      const int col = cb0 + jl, j = col % N, b = col / N;
      if (j >= from)
        for (int a = 0; a < 4; ++a) dg[16 * j + 4 * a + b] = T[(size_t)jl * n + j + a * N];
    }
    for (int q = 0; q < kLocalWindow * 4; ++q) {
      const int site = win0 + (q >> 2);
      if (site >= N) break;
      const int col = site + (q & 3) * N;
      if (col < cb0 || col >= cb1) continue;
      const cplx* src = T + (size_t)(col - cb0) * n;
      cplx* dst = wn + (size_t)q * n;
      for (int r = tid; r < n; r += blockDim.x) dst[r] = src[r];
    }
    ++rnd;
    if (tid == 0) pub_site = win0;
    group_sync(g);
    return !group_failed(g);
  };

  for (;;) {
    if (tid < 64) {
      const int lane = tid & 63, ea = (lane & 15) >> 2, eb = lane & 3;
      const cplx* dg = c.lu_diag + (size_t)(rnd & 1) * 16 * N;
      const bool from_g = rnd == 0;
      int st;
      const int site = evaluate_proposals(
          c, s, p, sh, box, dolog, w == 0,
          [&](int j) { return from_g ? G0[(size_t)(j + eb * N) * n + (j + ea * N)] : dg[16 * j + 4 * ea + eb]; },
          &st, N);
      if (tid == 0) {
        sh.site = site;
        sh.stop = st;
      }
    }
    __syncthreads();
    const int site = sh.site;
    stop = sh.stop;
    if (site >= N) break;
    // the site's four columns of the current G: global G before the first round, else the window
    if (rnd > 0 && !(site >= pub_site && site < pub_site + kLocalWindow)) {
      if (!publish_round(site, site)) return;
    }
    const cplx* colsrc[4];
    for (int k = 0; k < 4; ++k)
      colsrc[k] = rnd == 0 ? G0 + (size_t)(site + k * N) * n
                           : c.lu_win + (size_t)(rnd & 1) * win_sz + (size_t)((site - pub_site) * 4 + k) * n;
    // B = delta G[r_i, own columns] (4 x ncol), from the tile
    for (int jl = tid; jl < ncol; jl += blockDim.x) {
      cplx g4[4];
      for (int k = 0; k < 4; ++k) g4[k] = T[(size_t)jl * n + site + k * N];
      for (int k = 0; k < 4; ++k) {
        cplx acc = mk(0, 0);
        for (int m = 0; m < 4; ++m) cfma(acc, sh.delta[k * 4 + m], g4[m]);
        Brow[(size_t)k * ncol + jl] = acc;
      }
    }
    __syncthreads();
    // T[r, :] += A[r, :] B with A = (G[:, c_i] - E) M^-1 formed per row (update_greens!,
    // local_updates.jl:61-95)
    for (int r = tid; r < n; r += blockDim.x) {
      cplx g4[4], a4[4];
      for (int k = 0; k < 4; ++k) {
        g4[k] = colsrc[k][r];
        if (r == site + k * N) g4[k].re -= 1.0;
      }
      for (int k = 0; k < 4; ++k) {
        cplx acc = mk(0, 0);
        for (int m = 0; m < 4; ++m) cfma(acc, g4[m], sh.minv[m * 4 + k]);
        a4[k] = acc;
      }
      for (int jl = 0; jl < ncol; ++jl) {
        cplx acc = T[(size_t)jl * n + r];
        for (int k = 0; k < 4; ++k) cfma(acc, a4[k], Brow[(size_t)k * ncol + jl]);
        T[(size_t)jl * n + r] = acc;
      }
    }
    __syncthreads();
    if (!publish_round(site + 1, site + 1)) return;
  }
  cplx* Gout = c.mat[M_G];
  for (size_t e = tid; e < (size_t)ncol * n; e += blockDim.x) Gout[(size_t)cb0 * n + e] = T[e];
  if (w == 0) finish_slice(c, l, s, p, dolog, stop);
}

// ---------------------------------------------------------------------------
// Variant 3: delayed updates. Workgroup 0 runs the proposals alone and keeps every
// accepted rank-4 update as a pair of factors, A_a = (G'[:, c] - E) M^-1 (n x 4) and
// B_a = delta G'[r, :] (4 x n), where G' = G + sum_{pending} A B is the current Green's
// function; the diagonal site blocks of the sites still to visit are kept current in LDS
// (one 4x4 correction per accepted update and remaining site), so decisions read only LDS.
// After kLocalDelay pending updates (and at the end of the slice) every workgroup of the
// chain adds sum_a A_a B_a to its own columns of G: two group barriers per flush instead
// of one per accepted proposal, and G is streamed once per flush.
// Same arithmetic as update_greens! (local_updates.jl:61-95) up to the order of sums.
// ---------------------------------------------------------------------------
constexpr int kLocalDelay = 8;
enum : int { LD_CONTINUE = 0, LD_FLUSH = 1, LD_DONE = 2, LD_FLUSH_DONE = 3 };

__host__ __device__ inline size_t delayed_lds_bytes(int n, int N) {
  return slice_lds_doubles(N) * 8 + (size_t)N * 16 * 16 + (size_t)8 * n * 16;
}

// G[:, cb0:cb1] += sum_{a < m} A_a B_a[:, cb0:cb1], B staged through LDS in column chunks
__device__ inline void delayed_flush(const ChainDesc& c, int m, int cb0, int cb1, cplx* stage, int stage_cols) {
  const int n = c.n, tid = threadIdx.x;
  cplx* G = c.mat[M_G];
  const cplx* As = c.ld_A;
  const cplx* Bs = c.ld_B;
  const int kq = 4 * m;
  for (int j0 = cb0; j0 < cb1; j0 += stage_cols) {
    const int nc = min(stage_cols, cb1 - j0);
    __syncthreads();
    for (int e = tid; e < kq * nc; e += blockDim.x) {
      const int aq = e / nc, jl = e - aq * nc;
      stage[e] = Bs[(size_t)aq * n + j0 + jl];  // B_a[q][col], aq = 4 a + q
    }
    __syncthreads();
    for (int r = tid; r < n; r += blockDim.x) {
      for (int a0 = 0; a0 < m; a0 += 4) {
        cplx ar[16];
#pragma unroll
        for (int t = 0; t < 16; ++t) {
          const int a = a0 + (t >> 2);
          ar[t] = a < m ? As[((size_t)a * n + r) * 4 + (t & 3)] : mk(0, 0);
        }
        for (int jl = 0; jl < nc; ++jl) {
          cplx acc = mk(0, 0);
#pragma unroll
          for (int t = 0; t < 16; ++t)
            if (a0 + (t >> 2) < m) cfma(acc, ar[t], stage[(size_t)(4 * a0 + t) * nc + jl]);
          cplx* gp = G + (size_t)(j0 + jl) * n + r;
          *gp = *gp + acc;
        }
      }
    }
  }
  __syncthreads();
}

__global__ __launch_bounds__(256) void k_local_updates_delayed(const ChainDesc* chains, int slice, const int* slices,
                                                               double box, int dolog) {
  const ChainDesc& c = chains[blockIdx.y];
  const int n = c.n, N = c.N;
  const int W = gridDim.x, w = blockIdx.x, tid = threadIdx.x;
  const int l = slices ? slices[blockIdx.y] : slice;
  GroupSync g = chain_group(c);
  const int cb0 = (int)(((long long)n * w) / W), cb1 = (int)(((long long)n * (w + 1)) / W);

  extern __shared__ double smem[];
  cplx* Gd = reinterpret_cast<cplx*>(smem + slice_lds_doubles(N));  // [N][16] current diagonal blocks
  cplx* Am = Gd + (size_t)N * 16;                                   // [n][4] A of the last accepted
  cplx* Bm = Am + (size_t)4 * n;                                    // [4][n] B of the last accepted
  cplx* stage = Am;                                                  // flush staging (8n entries)
  const int stage_cols = (8 * n) / (4 * kLocalDelay);
  __shared__ LocalShared sh;
  __shared__ cplx bcol[kLocalDelay * 16];  // B_a[:, c_i] of the pending updates
  __shared__ cplx arow[kLocalDelay * 16];  // A_a[r_i, :] of the pending updates
  __shared__ int s_cmd;
  cplx* G = c.mat[M_G];

  if (w != 0) {
    // flush helper: wait for workgroup 0's command at each round
    for (;;) {
      group_sync(g);
      if (group_failed(g)) return;
      const int cmd = *reinterpret_cast<volatile int*>(c.ld_cmd);
      const int m = *reinterpret_cast<volatile int*>(c.ld_cmd + 1);
      if (cmd == LD_FLUSH || cmd == LD_FLUSH_DONE) {
        delayed_flush(c, m, cb0, cb1, stage, stage_cols);
        group_sync(g);
        if (group_failed(g)) return;
      }
      if (cmd == LD_DONE || cmd == LD_FLUSH_DONE) return;
    }
  }

  const SliceLds s = load_slice(c, l, smem);
  for (int e = tid; e < N * 16; e += blockDim.x) {
    const int j = e >> 4, a = (e >> 2) & 3, b = e & 3;
    Gd[e] = G[(size_t)(j + b * N) * n + (j + a * N)];
  }
  __syncthreads();
  Proposals p{s.cur0, dolog ? *c.log_count : 0, 0.0, c.scal[S_MINRATIO], 0, 0, 0, -1, 0.0, 0.0, 0.0, 0.0};
  int m = 0, stop = 0;
  cplx* As = c.ld_A;
  cplx* Bs = c.ld_B;

  auto round = [&](int cmd) -> bool {
    if (tid == 0) {
      c.ld_cmd[0] = cmd;
      c.ld_cmd[1] = m;
    }
    group_sync(g);
    if (group_failed(g)) return false;
    if (cmd == LD_FLUSH || cmd == LD_FLUSH_DONE) {
      delayed_flush(c, m, cb0, cb1, stage, stage_cols);
      group_sync(g);
      if (group_failed(g)) return false;
    }
    return true;
  };

  for (;;) {
    if (tid < 64) {
      const int lane = tid & 63, ea = (lane & 15) >> 2, eb = lane & 3;
      int st;
      const int site = evaluate_proposals(c, s, p, sh, box, dolog, true,
                                          [&](int j) { return Gd[16 * j + 4 * ea + eb]; }, &st, N);
      if (tid == 0) {
        sh.site = site;
        sh.stop = st;
      }
    }
    __syncthreads();
    const int site = sh.site;
    stop = sh.stop;
    if (site >= N) break;
    // corrections of the pending updates at the site's columns and rows
    for (int e = tid; e < m * 16; e += blockDim.x) {
      const int a = e >> 4, q = (e >> 2) & 3, k = e & 3;
      bcol[e] = Bs[((size_t)a * 4 + q) * n + site + k * N];        // B_a[q][c_k]
      arow[e] = As[((size_t)a * n + site + q * N) * 4 + k];        // A_a[r_q][k]
    }
    __syncthreads();
    // A = (G'[:, c] - E) M^-1, thread per row; every load of a row is issued before use
    for (int r = tid; r < n; r += blockDim.x) {
      cplx g4[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) g4[k] = G[(size_t)(site + k * N) * n + r];
      for (int a0 = 0; a0 < m; a0 += 4) {
        cplx av[16];
#pragma unroll
        for (int t = 0; t < 16; ++t)
          av[t] = a0 + (t >> 2) < m ? As[((size_t)(a0 + (t >> 2)) * n + r) * 4 + (t & 3)] : mk(0, 0);
#pragma unroll
        for (int da = 0; da < 4; ++da) {
          const int a = a0 + da;
          if (a >= m) break;
#pragma unroll
          for (int k = 0; k < 4; ++k) {
            const cplx* bc = bcol + a * 16 + k;
            cfma(g4[k], av[4 * da + 0], bc[0]);
            cfma(g4[k], av[4 * da + 1], bc[4]);
            cfma(g4[k], av[4 * da + 2], bc[8]);
            cfma(g4[k], av[4 * da + 3], bc[12]);
          }
        }
      }
      for (int k = 0; k < 4; ++k)
        if (r == site + k * N) g4[k].re -= 1.0;
      for (int k = 0; k < 4; ++k) {
        cplx acc = mk(0, 0);
        for (int q = 0; q < 4; ++q) cfma(acc, g4[q], sh.minv[q * 4 + k]);
        Am[4 * (size_t)r + k] = acc;
        As[((size_t)m * n + r) * 4 + k] = acc;
      }
    }
    // B = delta G'[r, :], thread per column; loads issued together as above
    for (int col = tid; col < n; col += blockDim.x) {
      cplx g4[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) g4[q] = G[(size_t)col * n + site + q * N];
      for (int a0 = 0; a0 < m; a0 += 4) {
        cplx bv[16];
#pragma unroll
        for (int t = 0; t < 16; ++t) bv[t] = a0 + (t >> 2) < m ? Bs[(size_t)(4 * a0 + t) * n + col] : mk(0, 0);
#pragma unroll
        for (int da = 0; da < 4; ++da) {
          const int a = a0 + da;
          if (a >= m) break;
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            const cplx* ar = arow + a * 16 + q * 4;
            cfma(g4[q], ar[0], bv[4 * da + 0]);
            cfma(g4[q], ar[1], bv[4 * da + 1]);
            cfma(g4[q], ar[2], bv[4 * da + 2]);
            cfma(g4[q], ar[3], bv[4 * da + 3]);
          }
        }
      }
      for (int k = 0; k < 4; ++k) {
        cplx acc = mk(0, 0);
        for (int q = 0; q < 4; ++q) cfma(acc, sh.delta[k * 4 + q], g4[q]);
        Bm[(size_t)k * n + col] = acc;
        Bs[((size_t)m * 4 + k) * n + col] = acc;
      }
    }
    __syncthreads();
    // diagonal blocks of the sites still to visit: Gd_j += A[r_j, :] B[:, c_j]
    for (int e = tid; e < (N - site - 1) * 16; e += blockDim.x) {
      const int j = site + 1 + (e >> 4), a = (e >> 2) & 3, b = e & 3;
      const cplx* ar = Am + 4 * (size_t)(j + a * N);
      cplx acc = Gd[16 * j + 4 * a + b];
      for (int q = 0; q < 4; ++q) cfma(acc, ar[q], Bm[(size_t)q * n + j + b * N]);
      Gd[16 * j + 4 * a + b] = acc;
    }
    ++m;
    __syncthreads();
    if (m == kLocalDelay) {
      if (!round(LD_FLUSH)) return;
      m = 0;
    }
  }
  if (!round(m > 0 ? LD_FLUSH_DONE : LD_DONE)) return;
  finish_slice(c, l, s, p, dolog, stop);
}

// ---------------------------------------------------------------------------
// Variant 4: windowed delayed updates. The sites of the slice are split into R windows
// (R even, at most K sites each). Inside a window no workgroup communicates:
//  * the accepted updates are kept implicitly. With G0 the window's starting G and
//    Yt = [G0[:, c_b] - E_b]_b the starting site columns of the accepted sites (minus their
//    unit columns), every factor A_b = (G'[:, c_b] - E_b) M_b^-1 equals Yt P_b for a small
//    coefficient matrix P_b (4K x 4): P_b = (S_b + sum_{b'<b} P_b' B_b'[:, c_b]) M_b^-1, S_b
//    selecting block b. Only P_b, the 4x4 blocks A_b(r_j, :) and B_b(:, c_j) at the window's
//    sites, B_b on the workgroup's own columns, and the current blocks G'(r_j, c_k) between
//    the window's sites are stored (LDS), so a window's size is not bounded by n x 4 factors;
//  * decisions read G'(r_j, c_j); new factor pairs need only the stored blocks and G0 on
//    the own columns.
// At the window's end each workgroup writes its columns of G0 + Yt W, W = sum_b P_b B_b
// (formed in place of the B_b), into the partner buffer (M_G and M_GALT alternate; R is
// even, so the slice ends in M_G), then one group barrier. Global G is read only from the
// current window's buffer, which no workgroup writes during the window. With TILE the own
// columns also stay in LDS for the slice.
// Same arithmetic as update_greens! (local_updates.jl:61-95) up to the order of sums.
// ---------------------------------------------------------------------------
constexpr int kWinMax = 12;     // window sites with LDS-resident own columns
constexpr int kWinMaxGlob = 12;  // window sites otherwise (Bo and H of the own columns bound it)
constexpr size_t kWindowLdsMax = 160 * 1024 - 1024;  // the CU's LDS less this kernel's static part

__host__ __device__ inline int window_bo_stride(int n, int wg) { return (n + wg - 1) / wg + 1; }
__host__ __device__ inline size_t window_lds_bytes(int n, int N, int wg, int K, bool tile) {
  const size_t ncs = window_bo_stride(n, wg);
  const size_t t = tile ? (size_t)(ncs - 1) * n : 0;
  return slice_lds_doubles(N) * 8 + (t + (size_t)8 * K * ncs + (size_t)48 * K * K + (size_t)32 * K) * 16;
}

// End of a window without LDS-resident columns, on the FP64 matrix cores:
// Gn[:, own] = Gc[:, own] + Yt W as transposed 16 x 16 tiles, D = W^T Yt^T + Gc^T, so that a
// tile's lane (i, j) holds own column i, row j and every load and store of G covers 16
// consecutive rows of a column. Each wave takes row blocks of 16; for its block the Yt^T
// fragments of all accepted sites stay in registers (two doubles per site and lane) while it
// walks the own columns two tiles at a time (four independent accumulator chains; the next
// pair's G loads are issued before this pair's products). One complex k-step of four
// (one accepted site's four columns of Yt) is four real MFMAs:
// re += Wr Yr - Wi Yi, im += Wr Yi + Wi Yr. v_mfma_f64_16x16x4 lane layout: A(i, k) and
// B(k, j) on lane i + 16 k (resp. j + 16 k), D(i, j) element v on lane j + 16 ((i - v * 4)):
// i = lane / 16 + 4 v. m >= 1.
typedef double v4d __attribute__((ext_vector_type(4)));
__device__ inline void window_end_mfma(const gcplx* __restrict__ Gc, gcplx* __restrict__ Gn, const cplx* W,
                                       const int* accs, int m, int n, int N, int cb0, int ncol, int ncs) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
  const int li = lane & 15, lk = lane >> 4;
  const int nrb = (n + 15) >> 4, ncp = (ncol + 31) >> 5;  // row blocks, pairs of column tiles
  for (int rb = wave; rb < nrb; rb += nw) {
    const int row = rb * 16 + li, rr = min(row, n - 1);
    // G tile loads of a pair of column tiles starting at own column c0
    auto load_pair = [&](int c0, cplx (&g)[8]) __attribute__((always_inline)) {
#pragma unroll
      for (int v = 0; v < 8; ++v) {
        const int col = min(c0 + lk + 4 * (v & 3) + 16 * (v >> 2), ncol - 1);
        g[v] = ld(Gc + (size_t)(cb0 + col) * n + rr);
      }
    };
    // the first two pairs' loads go out before the Yt fragments', and every pair's loads two
    // pairs ahead of its products
    cplx nxt[8], nx2[8];
    load_pair(0, nxt);
    load_pair(min(32, (ncp - 1) * 32), nx2);
    double yr[kWinMaxGlob], yi[kWinMaxGlob];
#pragma unroll
    for (int b = 0; b < kWinMaxGlob; ++b) {
      yr[b] = 0.0;
      yi[b] = 0.0;
      if (b < m) {  // wave-uniform
        const int col = accs[b] + lk * N;
        cplx v = ld(Gc + (size_t)col * n + rr);
        if (rr == col) v.re -= 1.0;
        yr[b] = v.re;
        yi[b] = v.im;
      }
    }
    for (int cp = 0; cp < ncp; ++cp) {
      const int c0 = cp * 32;
      v4d ar0, ai0, ar1, ai1;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        ar0[v] = nxt[v].re;
        ai0[v] = nxt[v].im;
        ar1[v] = nxt[v + 4].re;
        ai1[v] = nxt[v + 4].im;
      }
#pragma unroll
      for (int v = 0; v < 8; ++v) nxt[v] = nx2[v];
      if (cp + 2 < ncp) load_pair(c0 + 64, nx2);
      const int w0 = min(c0 + li, ncol - 1), w1 = min(c0 + 16 + li, ncol - 1);
#pragma unroll
      for (int b = 0; b < kWinMaxGlob; ++b) {
        if (b < m) {
          const cplx* wr = W + (size_t)(4 * b + lk) * ncs;
          const cplx x0 = wr[w0], x1 = wr[w1];
          ar0 = __builtin_amdgcn_mfma_f64_16x16x4f64(x0.re, yr[b], ar0, 0, 0, 0);
          ai0 = __builtin_amdgcn_mfma_f64_16x16x4f64(x0.re, yi[b], ai0, 0, 0, 0);
          ar1 = __builtin_amdgcn_mfma_f64_16x16x4f64(x1.re, yr[b], ar1, 0, 0, 0);
          ai1 = __builtin_amdgcn_mfma_f64_16x16x4f64(x1.re, yi[b], ai1, 0, 0, 0);
          ar0 = __builtin_amdgcn_mfma_f64_16x16x4f64(-x0.im, yi[b], ar0, 0, 0, 0);
          ai0 = __builtin_amdgcn_mfma_f64_16x16x4f64(x0.im, yr[b], ai0, 0, 0, 0);
          ar1 = __builtin_amdgcn_mfma_f64_16x16x4f64(-x1.im, yi[b], ar1, 0, 0, 0);
          ai1 = __builtin_amdgcn_mfma_f64_16x16x4f64(x1.im, yr[b], ai1, 0, 0, 0);
        }
      }
      if (row < n) {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int ca = c0 + lk + 4 * v, cb = ca + 16;
          if (ca < ncol) st_nt(Gn + (size_t)(cb0 + ca) * n + row, mk(ar0[v], ai0[v]));
          if (cb < ncol) st_nt(Gn + (size_t)(cb0 + cb) * n + row, mk(ar1[v], ai1[v]));
        }
      }
    }
  }
}

template <bool TILE>
__global__ __launch_bounds__(256) void k_local_updates_window(const ChainDesc* chains, int slice, const int* slices,
                                                              double box, int dolog, int K, int lds_cap,
                                                              const int* bmap) {
  // grid (W, chains), or with a block map a 1-D grid where block b serves chain bmap[b] & 0xFFFF
  // as workgroup (bmap[b] >> 16) & 0xFF of (bmap[b] >> 24) (workgroup counts per chain)
  const int bm = bmap ? bmap[blockIdx.x] : 0;
  const int ci = bmap ? (bm & 0xFFFF) : (int)blockIdx.y;
  const ChainDesc& c = chains[ci];
  const int n = c.n, N = c.N;
  const int W = bmap ? (bm >> 24) & 0xFF : (int)gridDim.x, w = bmap ? (bm >> 16) & 0xFF : (int)blockIdx.x;
  const int tid = threadIdx.x;
  const int l = slices ? slices[ci] : slice;
  const GroupSync g{c.bar_count, c.bar_gen, c.error, (unsigned int)W};
  const int cb0 = (int)(((long long)n * w) / W), cb1 = (int)(((long long)n * (w + 1)) / W);
  const int ncol = cb1 - cb0;
  const int ncs = window_bo_stride(n, W);
  // K fits the largest chain of the launch; a smaller chain takes the largest window its own
  // layout fits in the allocated LDS (fewer windows, fewer passes of G)
  if (!TILE)
    for (int k2 = kWinMaxGlob; k2 > K; --k2)
      if (window_lds_bytes(n, N, W, k2, false) <= (size_t)lds_cap) {
        K = k2;
        break;
      }
  const int K4 = 4 * K;

  extern __shared__ double smem[];
  const SliceLds s = load_slice(c, l, smem);
  cplx* T = reinterpret_cast<cplx*>(smem + slice_lds_doubles(N));  // [ncs - 1][n] own columns (TILE)
  cplx* Bo = T + (TILE ? (size_t)(ncs - 1) * n : 0);  // [K][4][ncs] B_b on own columns
  // [K][4][ncs] H_j = G'(r_j, own) = G0(r_j, own) + sum_b A_b(r_j, :) B_b(:, own); W at a window's end
  cplx* Hs = Bo + (size_t)K4 * ncs;
  cplx* Gp = Hs + (size_t)K4 * ncs;  // [K][K][16]  G'(r_j, c_k)
  cplx* Pc = Gp + (size_t)16 * K * K;  // [K][4K][4] P_b
  cplx* Us = Pc + (size_t)16 * K * K;  // [K][4K][4] U_k = sum_b P_b B_b(:, c_k)
  cplx* Bw = Us + (size_t)16 * K * K;  // [K][16]    B_m(:, c_k) of the update being applied
  cplx* Ar = Bw + (size_t)16 * K;      // [K][16]    A_m(r_j, :)
  __shared__ LocalShared sh;
  __shared__ int accs[kWinMax];  // sites accepted in the current window, in order
  gcplx* X[2] = {gp(c.mat[M_G]), gp(c.mat[M_GALT])};
  int R = (N + K - 1) / K;
  R += R & 1;
  if (TILE) {
    const gcplx* G0 = X[0] + (size_t)cb0 * n;
    const int tot = ncol * n;
    for (int e0 = tid; e0 < tot; e0 += 4 * blockDim.x) {
      cplx v[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) v[q] = ld(G0 + min(e0 + q * (int)blockDim.x, tot - 1));
#pragma unroll
      for (int q = 0; q < 4; ++q)
        if (e0 + q * (int)blockDim.x < tot) {
          T[e0 + q * blockDim.x].re = v[q].re;  // (whole-struct copies from a register array
          T[e0 + q * blockDim.x].im = v[q].im;  // are lowered through a stack slot)
        }
    }
    __syncthreads();  // the first window's H rows are read from T
  }

  Proposals p{s.cur0, dolog ? *c.log_count : 0, 0.0, c.scal[S_MINRATIO], 0, 0, 0, -1, 0.0, 0.0, 0.0, 0.0};
  int stop = 0;  // the uniform stream ran out: later windows make no proposals
  for (int rw = 0; rw < R; ++rw) {
    const int s0 = (int)((long long)rw * N / R), s1 = (int)((long long)(rw + 1) * N / R), kw = s1 - s0;
    const gcplx* Gc = X[rw & 1];
    gcplx* Gn = X[(rw + 1) & 1];
    // Gp[j][k](a, b) = G(r_j + a N, c_k + b N) for the window's sites j, k; U = 0, H = 0
    // eight loads per thread in flight before the first store (one HBM latency per 2048
    // entries); the variant without LDS-resident columns has no registers to spare for it
    if (!TILE)
      for (int e = tid; e < kw * kw * 16; e += blockDim.x) {
        const int jk = e >> 4, jj = jk / kw, kk = jk - jj * kw, a = (e >> 2) & 3, b = e & 3;
        Gp[(jj * K + kk) * 16 + a * 4 + b] = ld(Gc + (size_t)(s0 + kk + b * N) * n + (s0 + jj + a * N));
      }
    for (int e0 = 0; TILE && e0 < kw * kw * 16; e0 += 8 * 256) {
      cplx v[8];
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const int e = min(e0 + tid + 256 * j, kw * kw * 16 - 1);
        const int jk = e >> 4, jj = jk / kw, kk = jk - jj * kw, a = (e >> 2) & 3, b = e & 3;
        v[j] = ld(Gc + (size_t)(s0 + kk + b * N) * n + (s0 + jj + a * N));
      }
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const int e = e0 + tid + 256 * j;
        if (e < kw * kw * 16) {
          const int jk = e >> 4, jj = jk / kw, kk = jk - jj * kw, a = (e >> 2) & 3, b = e & 3;
          Gp[(jj * K + kk) * 16 + a * 4 + b] = v[j];
        }
      }
    }
    for (int e = tid; e < kw * K4 * 4; e += blockDim.x) Us[e] = mk(0, 0);
    // H_j starts as the window's row r_j of G0 on the own columns (so an acceptance reads its
    // site's rows from LDS); four loads per thread in flight
    for (int e0 = tid; e0 < kw * 4 * ncs; e0 += 4 * blockDim.x) {
      cplx v[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int e = min(e0 + u * (int)blockDim.x, kw * 4 * ncs - 1);
        const int jq = e / ncs, jl = min(e - jq * ncs, ncol - 1), row = s0 + (jq >> 2) + (jq & 3) * N;
        v[u] = TILE ? T[(size_t)jl * n + row] : ld(Gc + (size_t)(cb0 + jl) * n + row);
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int e = e0 + u * (int)blockDim.x;
        if (e < kw * 4 * ncs) {
          Hs[e].re = v[u].re;
          Hs[e].im = v[u].im;
        }
      }
    }
    __syncthreads();
    // Per accepted proposal m at window site ii, two barriers: after the evaluation (S1) and
    // after the factor pair (S2). The sums kept for the sites ahead are then brought up to date
    // while wave 0 already evaluates the next proposals: wave 0 updates the diagonal blocks
    // G'(r_j, c_j) it reads, waves 1-3 everything else (off-diagonal blocks, U, H); the next
    // S1 completes both before the factor pair of the next acceptance reads them.
    int m = 0, pend = -1;  // pend: window site of the acceptance whose sums are still to update
    while (!stop) {
      if (tid < 64) {
        const int lane = tid & 63, ea = (lane & 15) >> 2, eb = lane & 3;
        if (pend >= 0) {  // G'(r_j, c_j) += A_m(r_j, :) B_m(:, c_j), j > pend
          const int ka = kw - pend - 1;
          for (int e = lane; e < ka * 16; e += 64) {
            const int jj = pend + 1 + (e >> 4), a = (e >> 2) & 3, b = e & 3;
            cplx acc = Gp[(jj * K + jj) * 16 + a * 4 + b];
#pragma unroll
            for (int q = 0; q < 4; ++q) cfma(acc, Ar[jj * 16 + a * 4 + q], Bw[jj * 16 + q * 4 + b]);
            Gp[(jj * K + jj) * 16 + a * 4 + b] = acc;
          }
          __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");  // the wave's writes before its reads
        }
        int st;
        const int site = evaluate_proposals(
            c, s, p, sh, box, dolog, w == 0,
            [&](int j) { return Gp[((j - s0) * (K + 1)) * 16 + ea * 4 + eb]; }, &st, s1, true);
        if (tid == 0) {
          sh.site = site;
          sh.stop = st;
        }
      } else if (pend >= 0) {
        // the pending acceptance (index m - 1) for the sites ahead (j, k > pend):
        //   G'(r_j, c_k) += A(r_j, :) B(:, c_k) (j != k),  U_k += P B(:, c_k),  H_j += A(r_j, :) B(:, own)
        const int t3 = tid - 64, n3 = blockDim.x - 64, mp = m - 1;
        const int ka = kw - pend - 1;
        for (int e = t3; e < ka * ka * 16; e += n3) {
          const int jk = e >> 4, jj = pend + 1 + jk / ka, kk = pend + 1 + jk % ka, a = (e >> 2) & 3, b = e & 3;
          if (jj == kk) continue;
          cplx acc = Gp[(jj * K + kk) * 16 + a * 4 + b];
#pragma unroll
          for (int q = 0; q < 4; ++q) cfma(acc, Ar[jj * 16 + a * 4 + q], Bw[kk * 16 + q * 4 + b]);
          Gp[(jj * K + kk) * 16 + a * 4 + b] = acc;
        }
        const int nu = 16 * m;
        for (int e = t3; e < ka * nu; e += n3) {
          const int kk = pend + 1 + e / nu, r2 = e % nu, t = r2 >> 2, q = r2 & 3;
          cplx acc = Us[((size_t)kk * K4 + t) * 4 + q];
#pragma unroll
          for (int q2 = 0; q2 < 4; ++q2) cfma(acc, Pc[((size_t)mp * K4 + t) * 4 + q2], Bw[kk * 16 + q2 * 4 + q]);
          Us[((size_t)kk * K4 + t) * 4 + q] = acc;
        }
        const int nh = 4 * ncol;
        for (int e = t3; e < ka * nh; e += n3) {
          const int jj = pend + 1 + e / nh, r2 = e % nh, q = r2 / ncol, jl = r2 - q * ncol;
          cplx acc = Hs[((size_t)jj * 4 + q) * ncs + jl];
#pragma unroll
          for (int k = 0; k < 4; ++k) cfma(acc, Ar[jj * 16 + q * 4 + k], Bo[((size_t)mp * 4 + k) * ncs + jl]);
          Hs[((size_t)jj * 4 + q) * ncs + jl] = acc;
        }
      }
      __syncthreads();  // S1
      const int site = sh.site;
      stop = sh.stop;
      if (site >= s1) break;
      const int ii = site - s0;
      if (tid < 64) {
        // wave 0: the accepted proposal's field update, while waves 1-3 form its factor pair
        commit_accept(c, s, p);
      } else {
        const int t3 = tid - 64, n3 = blockDim.x - 64;
        // B_m on own columns: delta G'(r_i, own) = delta H_i
        for (int jl = t3; jl < ncol; jl += n3) {
          cplx g4[4];
#pragma unroll
          for (int q = 0; q < 4; ++q) g4[q] = Hs[((size_t)ii * 4 + q) * ncs + jl];
#pragma unroll
          for (int k = 0; k < 4; ++k) {
            cplx acc = mk(0, 0);
#pragma unroll
            for (int q = 0; q < 4; ++q) cfma(acc, sh.delta[k * 4 + q], g4[q]);
            Bo[((size_t)m * 4 + k) * ncs + jl] = acc;
          }
        }
        // B_m(:, c_k) = delta G'(r_i, c_k) and A_m(r_j, :) = G'(r_j, c_i) M^-1 for the sites
        // ahead (j, k > i: the only ones read again)
        for (int e = t3; e < kw * 16; e += n3) {
          const int kk = e >> 4, k = (e >> 2) & 3, b = e & 3;
          if (kk <= ii) continue;
          cplx bw = mk(0, 0), ar = mk(0, 0);
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            cfma(bw, sh.delta[k * 4 + q], Gp[(ii * K + kk) * 16 + q * 4 + b]);
            cfma(ar, Gp[(kk * K + ii) * 16 + k * 4 + q], sh.minv[q * 4 + b]);
          }
          Bw[kk * 16 + k * 4 + b] = bw;
          Ar[kk * 16 + k * 4 + b] = ar;
        }
        // P_m = (S_m + U_i) M^-1 (rows of blocks <= m)
        for (int e = t3; e < 16 * (m + 1); e += n3) {
          const int t = e >> 2, k = e & 3;
          cplx acc = mk(0, 0);
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            cplx v = Us[((size_t)ii * K4 + t) * 4 + q];
            if (t == 4 * m + q) v.re += 1.0;
            cfma(acc, v, sh.minv[q * 4 + k]);
          }
          Pc[((size_t)m * K4 + t) * 4 + k] = acc;
        }
        if (t3 == 0) accs[m] = site;
      }
      __syncthreads();  // S2
      pend = ii;
      ++m;
    }
    // W = sum_b P_b B_b (into the H storage, no longer needed): entry (t, jl) sums over the b
    // with P_b nonzero in row t (b >= t / 4)
    cplx* Wb = Hs;
    for (int e = tid; e < 4 * m * ncol; e += blockDim.x) {
      const int t = e / ncol, jl = e - t * ncol;
      cplx acc = mk(0, 0);
      for (int b = t >> 2; b < m; ++b) {
        const cplx* pb = Pc + ((size_t)b * K4 + t) * 4;
        const cplx* bo = Bo + (size_t)b * 4 * ncs + jl;
#pragma unroll
        for (int q = 0; q < 4; ++q) cfma(acc, pb[q], bo[(size_t)q * ncs]);
      }
      Wb[(size_t)t * ncs + jl] = acc;
    }
    __syncthreads();
    // own columns of the partner buffer: G0 + Yt W (a plain copy when m = 0)
    if (TILE) {
      for (int r = tid; r < n; r += blockDim.x)
        for (int j0 = 0; j0 < ncol; j0 += 4) {
          cplx acc[4];
#pragma unroll
          for (int q = 0; q < 4; ++q) acc[q] = T[(size_t)min(j0 + q, ncol - 1) * n + r];
          // Yt's row r for accepted site b + 1 is loaded while site b's products are formed
          cplx yn[4];
          const int a0 = m > 0 ? accs[0] : 0;  // (accs[0] is unset in a window without acceptances)
#pragma unroll
          for (int k = 0; k < 4; ++k) yn[k] = ld(Gc + (size_t)(a0 + k * N) * n + r);
          for (int b = 0; b < m; ++b) {
            cplx y[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
              y[k] = yn[k];
              if (r == accs[b] + k * N) y[k].re -= 1.0;
            }
            const int bn = min(b + 1, m - 1);
#pragma unroll
            for (int k = 0; k < 4; ++k) yn[k] = ld(Gc + (size_t)(accs[bn] + k * N) * n + r);
#pragma unroll
            for (int k = 0; k < 4; ++k) {
              const cplx* wr = Wb + ((size_t)b * 4 + k) * ncs;
#pragma unroll
              for (int q = 0; q < 4; ++q) cfma(acc[q], y[k], wr[min(j0 + q, ncol - 1)]);
            }
          }
#pragma unroll
          for (int q = 0; q < 4; ++q)
            if (j0 + q < ncol) {
              T[(size_t)(j0 + q) * n + r].re = acc[q].re;
              T[(size_t)(j0 + q) * n + r].im = acc[q].im;
              st_nt(Gn + (size_t)(cb0 + j0 + q) * n + r, acc[q]);
            }
        }
    } else if (m == 0) {
      for (int e = tid; e < ncol * n; e += blockDim.x) {
        st_nt(Gn + (size_t)cb0 * n + e, ld(Gc + (size_t)cb0 * n + e));
      }
    } else {
      window_end_mfma(Gc, Gn, Wb, accs, m, n, N, cb0, ncol, ncs);
    }
    if (rw + 1 < R) {
      group_sync(g);
      if (group_failed(g)) return;
    }
  }
  if (w == 0) finish_slice(c, l, s, p, dolog, stop);
}

// the window of variant 4 at this shape: its K, and whether the own columns fit in LDS too
// (preferred when the window keeps at least 4 sites); K = 0 when nothing fits
int local_window_sites(int nmax, int wg, bool* tile) {
  for (int K = kWinMax; K >= 4; --K)
    if (window_lds_bytes(nmax, nmax / 4, wg, K, true) <= kWindowLdsMax) {
      if (tile) *tile = true;
      return K;
    }
  if (tile) *tile = false;
  for (int K = kWinMaxGlob; K >= 2; --K)
    if (window_lds_bytes(nmax, nmax / 4, wg, K, false) <= kWindowLdsMax) return K;
  return 0;
}

bool local_delayed_fits(int nmax) { return delayed_lds_bytes(nmax, nmax / 4) <= 150 * 1024; }

// LDS bytes of the two variants at the given shape (dynamic part)
static size_t local_lds_global(int n, int N, int wg) {
  return slice_lds_doubles(N) * 8 + (size_t)4 * n * 16 + (size_t)4 * ((n + wg - 1) / wg + 1) * 16;
}
static size_t local_lds_tile(int n, int N, int wg) {
  const size_t ncol = (size_t)(n + wg - 1) / wg;
  return slice_lds_doubles(N) * 8 + ncol * n * 16 + (size_t)4 * ncol * 16;
}
constexpr size_t kLocalLdsMax = 150 * 1024;

bool local_tile_fits(int nmax, int wg) { return local_lds_tile(nmax, nmax / 4, wg) <= kLocalLdsMax; }

// Workgroups per chain of a window-variant launch without LDS-resident columns over chains of
// sizes n[k] (a block map for launch_local_updates): every chain starts at the fewest
// workgroups whose layout holds a window of 4 sites, then the chain whose workgroups each move
// the most G per slice -- (windows per slice) x n^2 / W -- takes one more, until 256 (one per
// CU) are given out; at most n / 4 and 255 per chain. -1 when the chains do not fit.
int local_block_map(const std::vector<int>& n, std::vector<int>& map) {
  const int na = (int)n.size();
  auto kfit = [](int nn, int W) {
    for (int K = kWinMaxGlob; K >= 2; --K)
      if (window_lds_bytes(nn, nn / 4, W, K, false) <= kWindowLdsMax) return K;
    return 0;
  };
  auto load = [&](int nn, int W) {
    const int K = kfit(nn, W), N = nn / 4;
    int R = (N + K - 1) / K;
    R += R & 1;
    return (double)R * nn * nn / W;
  };
  std::vector<int> W(na);
  int sum = 0;
  for (int k = 0; k < na; ++k) {
    int w = 1;
    while (w < 255 && kfit(n[k], w) < 4) ++w;
    if (kfit(n[k], w) < 2) return -1;
    W[k] = w;
    sum += w;
  }
  if (sum > 256) return -1;
  while (sum < 256) {
    int best = -1;
    double bl = 0.0;
    for (int k = 0; k < na; ++k)
      if (W[k] < std::min(255, n[k] / 4)) {
        const double v = load(n[k], W[k]);
        if (v > bl) {
          bl = v;
          best = k;
        }
      }
    if (best < 0) break;
    ++W[best];
    ++sum;
  }
  map.clear();
  for (int k = 0; k < na; ++k)
    for (int w = 0; w < W[k]; ++w) map.push_back(k | (w << 16) | (W[k] << 24));
  return (int)map.size();
}

hipError_t launch_local_updates(hipStream_t s, const ChainDesc* chains, int nchains, int nmax, int slice,
                                const int* slices, double box, int wg, int log, int variant, const int* bmap,
                                int nblocks) {
  const int Nmax = nmax / 4;
  if (variant < 1 || variant > 4) return hipErrorInvalidValue;
  if (bmap && variant != 4) return hipErrorInvalidValue;
  bool tile = false;
  // with a block map every chain sizes its own window (no LDS-resident columns)
  int K = bmap ? 2 : variant == 4 ? local_window_sites(nmax, wg, &tile) : 0;
  if (variant == 4 && K < 2) return hipErrorInvalidValue;
  const void* fn = variant == 2   ? (const void*)k_local_updates_tile
                   : variant == 3 ? (const void*)k_local_updates_delayed
                   : variant == 4 ? (tile ? (const void*)k_local_updates_window<true>
                                          : (const void*)k_local_updates_window<false>)
                                  : (const void*)k_local_updates;
  // the window variant without LDS-resident columns takes the whole budget: its smaller
  // chains size their windows to it
  const size_t lds = variant == 2   ? local_lds_tile(nmax, Nmax, wg)
                     : variant == 3 ? delayed_lds_bytes(nmax, Nmax)
                     : variant == 4 ? (tile ? window_lds_bytes(nmax, Nmax, wg, K, tile) : kWindowLdsMax)
                                    : local_lds_global(nmax, Nmax, wg);
  if (lds > (variant == 4 ? kWindowLdsMax : kLocalLdsMax)) return hipErrorInvalidValue;
  {
    hipError_t e = ensure_dynamic_lds(fn, lds);
    if (e != hipSuccess) return e;
  }
  int lds_cap = (int)lds;
  void* args1[] = {(void*)&chains, (void*)&slice, (void*)&slices, (void*)&box, (void*)&log, (void*)&K};
  void* args4[] = {(void*)&chains, (void*)&slice, (void*)&slices,  (void*)&box,
                   (void*)&log,    (void*)&K,     (void*)&lds_cap, (void*)&bmap};
  const dim3 grid = bmap ? dim3(nblocks, 1) : dim3(wg, nchains);
  hipError_t e = hipLaunchCooperativeKernel(fn, grid, dim3(256), variant == 4 ? args4 : args1, lds, s);
  if (e != hipSuccess) (void)hipGetLastError();
  return e;
}

// ---------------------------------------------------------------------------
// Single-proposal surfaces: calc_detratio (local_updates.jl:42-59) and a forced
// update_greens! (local_updates.jl:61-95) with the stored delta and M^-1.
// out[0..1] = det, out[2..33] = delta (column-major 4x4, interleaved),
// out[34..65] = M (column-major 4x4, interleaved).
// ---------------------------------------------------------------------------
__global__ void k_detratio(const ChainDesc* chains, int slice, int site, const double* new_op, double* out) {
  const ChainDesc& c = chains[0];
  if (threadIdx.x != 0) return;
  const int N = c.N, n = c.n, i = site;
  const double* f = c.phi + 3 * ((size_t)i + (size_t)N * slice);
  cplx E1[16], E2[16], Dm[16], Mm[16], Gm[16];
  ev_op(c.lam * c.dtau, f[0], f[1], f[2], -1.0, E1);
  ev_op(c.lam * c.dtau, new_op[0], new_op[1], new_op[2], 1.0, E2);
  mul4(E1, E2, Dm);
  for (int a = 0; a < 4; ++a) Dm[a * 4 + a].re -= 1.0;
  const cplx* G = c.mat[M_G];
  for (int a = 0; a < 4; ++a)
    for (int b = 0; b < 4; ++b) {
      const cplx gv = G[(size_t)(i + b * N) * n + (i + a * N)];
      Gm[a * 4 + b] = mk((a == b ? 1.0 : 0.0) - gv.re, -gv.im);
    }
  mul4(Dm, Gm, Mm);
  for (int a = 0; a < 4; ++a) Mm[a * 4 + a].re += 1.0;
  const cplx det = det4(Mm);
  out[0] = det.re;
  out[1] = det.im;
  for (int a = 0; a < 4; ++a)
    for (int b = 0; b < 4; ++b) {
      out[2 + 2 * (b * 4 + a)] = Dm[a * 4 + b].re;
      out[3 + 2 * (b * 4 + a)] = Dm[a * 4 + b].im;
      out[34 + 2 * (b * 4 + a)] = Mm[a * 4 + b].re;
      out[35 + 2 * (b * 4 + a)] = Mm[a * 4 + b].im;
    }
  cplx Mi[16];
  inv4(Mm, Mi);
  for (int e = 0; e < 16; ++e) {
    c.lu_small[e] = Dm[e];
    c.lu_small[16 + e] = Mi[e];
  }
}

hipError_t launch_detratio(hipStream_t s, const ChainDesc* chains, int slice, int site, const double* new_op,
                           double* out) {
  hipLaunchKernelGGL(k_detratio, dim3(1), dim3(64), 0, s, chains, slice, site, new_op, out);
  return hipGetLastError();
}

// G_alt = G + (G[:, c_i] - E) M^-1 delta G[r_i, :], then copied back (one workgroup per column tile)
__global__ __launch_bounds__(256) void k_force_update(const ChainDesc* chains, int site, int phase) {
  const ChainDesc& c = chains[0];
  const int n = c.n, N = c.N, i = site;
  const cplx* G = c.mat[M_G];
  cplx* Gn = c.mat[M_GALT];
  const cplx* Dm = c.lu_small;
  const cplx* Mi = c.lu_small + 16;
  const size_t nn = (size_t)n * n;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < nn; e += (size_t)gridDim.x * blockDim.x) {
    if (phase == 1) {
      c.mat[M_G][e] = Gn[e];
      continue;
    }
    const int j = (int)(e / n), r = (int)(e - (size_t)j * n);
    cplx g4[4], a[4], b[4], gr[4];
    for (int k = 0; k < 4; ++k) {
      g4[k] = G[(size_t)(i + k * N) * n + r];
      if (r == i + k * N) g4[k].re -= 1.0;
      gr[k] = G[(size_t)j * n + (i + k * N)];
    }
    for (int k = 0; k < 4; ++k) {
      a[k] = mk(0, 0);
      b[k] = mk(0, 0);
      for (int m = 0; m < 4; ++m) {
        cfma(a[k], g4[m], Mi[m * 4 + k]);
        cfma(b[k], Dm[k * 4 + m], gr[m]);
      }
    }
    cplx acc = G[e];
    for (int k = 0; k < 4; ++k) cfma(acc, a[k], b[k]);
    Gn[e] = acc;
  }
}

hipError_t launch_force_update(hipStream_t s, const ChainDesc* chains, int nmax, int site) {
  int blocks = (nmax * nmax + 255) / 256;
  if (blocks > 1024) blocks = 1024;
  hipLaunchKernelGGL(k_force_update, dim3(blocks), dim3(256), 0, s, chains, site, 0);
  hipLaunchKernelGGL(k_force_update, dim3(blocks), dim3(256), 0, s, chains, site, 1);
  return hipGetLastError();
}

__global__ void k_reset_stats(const ChainDesc* chains) {
  const ChainDesc& c = chains[blockIdx.x];
  if (threadIdx.x == 0) {
    c.scal[S_ACC_SLICE] = 0;
    c.scal[S_ACC_TOTAL] = 0;
    c.scal[S_PROP_TOTAL] = 0;
    c.scal[S_DRIFT_LAST] = 0;
    c.scal[S_DRIFT_MAX] = 0;
    c.scal[S_NONREAL] = 0;
    c.scal[S_MINRATIO] = 1e300;
    c.scal[S_GLOBAL_NACC] = 0;
    c.scal[S_GLOBAL_NPROP] = 0;
    *c.log_count = 0;
  }
}
hipError_t launch_reset_stats(hipStream_t s, const ChainDesc* chains, int nchains) {
  hipLaunchKernelGGL(k_reset_stats, dim3(nchains), dim3(64), 0, s, chains);
  return hipGetLastError();
}

}  // namespace dqmc
