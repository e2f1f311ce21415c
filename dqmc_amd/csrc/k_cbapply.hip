// Slice-matrix application B_l = H_B^(1/2) H_A H_B^(1/2) e^{dtau mu} eV_l and its
// inverse / adjoint, from the left (on columns) or the right (on rows).
//
// Reference: slice_matrices.jl:101-226 (Assaad checkerboard, five sparse passes
// each followed by a copy), interactions.jl:35-88 (eV rebuilt per call) and
// slice_matrices.jl:25-94 (dense, no checkerboard). Here every factor is a set
// of disjoint 4x4 mixes (a site's 4 flavours for eV, a plaquette's 4 sites
// within one flavour block for the hoppings, a diagonal for mu), so a tile of
// vectors is loaded into LDS once, all five factors are applied in LDS, and the
// tile is written back: one HBM read and one write of the operand in total.
// The generic checkerboard (hoppings_checkerboard_generic.jl, slice_matrices.jl:233-393) has
// G groups of disjoint bonds instead of two groups of plaquettes; its hopping factors are 2x2
// mixes on a bond's two sites, B = H_G^(1/2) ... H_2^(1/2) H_1 H_2^(1/2) ... H_G^(1/2) mu eV,
// the same stage pattern with G groups.
#include "chain.hpp"

namespace dqmc {

enum StageKind : int { ST_EV = 0, ST_MU = 1, ST_MUINV = 2, ST_PLAQ = 3, ST_BOND = 4 };
enum Mode : int { MODE_N = 0, MODE_T = 1, MODE_C = 2 };
// plaquette variants in ChainDesc::plaq
enum PlaqVar : int { PV_HALF = 0, PV_HALF_INV = 1, PV_FULL = 2, PV_FULL_INV = 3 };

struct Stage {
  int kind, mode, var, group;
  double power;
};
struct StageList {
  Stage st[2 * kMaxBondGroups + 1];
  int count;
};

// eV site block (interactions.jl:102-140): [[C,S,0,R],[S*,C,-R,0],[0,-R,C,S*],[R,0,S,C]]
__device__ inline void ev_block(double C, cplx S, double R, cplx E[16]) {
  const cplx z = mk(0, 0);
  E[0] = mk(C, 0);   E[1] = S;          E[2] = z;          E[3] = mk(R, 0);
  E[4] = conj(S);    E[5] = mk(C, 0);   E[6] = mk(-R, 0);  E[7] = z;
  E[8] = z;          E[9] = mk(-R, 0);  E[10] = mk(C, 0);  E[11] = conj(S);
  E[12] = mk(R, 0);  E[13] = z;         E[14] = S;         E[15] = mk(C, 0);
}

// y = op(P) x for a 4x4 row-major P; the mode is a template argument so that the loops that
// call it hold no branch (stages dispatch on the mode once)
template <int MODE>
__device__ inline void mix4(const cplx* P, cplx x[4], cplx y[4]) {
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    cplx acc = mk(0, 0);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      if (MODE == MODE_N)
        cfma(acc, P[a * 4 + k], x[k]);
      else if (MODE == MODE_T)
        cfma(acc, P[k * 4 + a], x[k]);
      else
        cfma_cj(acc, P[k * 4 + a], x[k]);
    }
    y[a] = acc;
  }
}

// the eV stage and a plaquette stage on the tile, for one mode
template <int MODE>
__device__ inline void ev_stage(cplx* vec, const double* evc, double power, int n, int N, int nv) {
  for (int idx = threadIdx.x; idx < nv * N; idx += blockDim.x) {
    const int v = idx / N, i = idx - v * N;
    cplx E[16];
    ev_block(evc[4 * i], mk(power * evc[4 * i + 1], power * evc[4 * i + 2]), power * evc[4 * i + 3], E);
    cplx x[4], y[4];
    cplx* base = vec + (size_t)v * n + i;
#pragma unroll
    for (int a = 0; a < 4; ++a) x[a] = base[a * N];
    mix4<MODE>(E, x, y);
#pragma unroll
    for (int a = 0; a < 4; ++a) base[a * N] = y[a];
  }
}
// UNI: every plaquette of a block has the same factor, staged in LDS (Puni); otherwise each
// plaquette's factor is read from global memory (kept apart: a pointer that may point to
// either compiles to FLAT loads)
template <int MODE, bool UNI>
__device__ inline void plaq_stage(cplx* vec, const cplx* P0, const cplx* Puni, const int* sites, int n, int N,
                                  int nv) {
  const int nplaq = N / 4;
  for (int idx = threadIdx.x; idx < nv * 4 * nplaq; idx += blockDim.x) {
    const int v = idx / (4 * nplaq);
    const int rem = idx - v * 4 * nplaq;
    const int b = rem / nplaq, pq = rem - b * nplaq;
    cplx P[16];
    if (UNI) {
#pragma unroll
      for (int e = 0; e < 16; ++e) P[e] = Puni[b * 16 + e];
    } else {
      const gcplx* Pg = gp(P0) + ((size_t)b * nplaq + pq) * 16;
#pragma unroll
      for (int e = 0; e < 16; ++e) P[e] = ld(Pg + e);
    }
    const gint* st4 = gp(sites) + 4 * pq;
    cplx* base = vec + (size_t)v * n + b * N;
    cplx x[4], y[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) x[a] = base[st4[a]];
    mix4<MODE>(P, x, y);
#pragma unroll
    for (int a = 0; a < 4; ++a) base[st4[a]] = y[a];
  }
}

// y = op(P) x for a 2x2 row-major P (a bond's two sites)
template <int MODE>
__device__ inline void mix2(const cplx* P, cplx x[2], cplx y[2]) {
#pragma unroll
  for (int a = 0; a < 2; ++a) {
    cplx acc = mk(0, 0);
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      if (MODE == MODE_N)
        cfma(acc, P[a * 2 + k], x[k]);
      else if (MODE == MODE_T)
        cfma(acc, P[k * 2 + a], x[k]);
      else
        cfma_cj(acc, P[k * 2 + a], x[k]);
    }
    y[a] = acc;
  }
}

// one group of the generic checkerboard on the tile: bonds q0 .. q1-1, blocks B0 [q][4][2][2]
template <int MODE>
__device__ inline void bond_stage(cplx* vec, const cplx* B0, const int* bsites, int q0, int q1, int n, int N,
                                  int nv) {
  const int nq = q1 - q0;
  for (int idx = threadIdx.x; idx < nv * 4 * nq; idx += blockDim.x) {
    const int v = idx / (4 * nq);
    const int rem = idx - v * 4 * nq;
    const int b = rem / nq, q = q0 + (rem - b * nq);
    const gcplx* Pg = gp(B0) + ((size_t)q * 4 + b) * 4;
    cplx P[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) P[e] = ld(Pg + e);
    const int s0 = gp(bsites)[2 * q], s1 = gp(bsites)[2 * q + 1];
    cplx* base = vec + (size_t)v * n + b * N;
    cplx x[2] = {base[s0], base[s1]}, y[2];
    mix2<MODE>(P, x, y);
    base[s0] = y[0];
    base[s1] = y[1];
  }
}

// C, S, R per site for power +1 (interactions.jl:52-58); the power sign is applied later.
__device__ inline void ev_coeffs(const ChainDesc& c, int l, int i, double& C, cplx& S, double& R) {
  const double* f = c.phi + 3 * ((size_t)i + (size_t)c.N * l);
  const double nrm = sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
  const double x = c.lam * c.dtau * nrm;
  const double sh = sinh(x) / nrm;
  C = cosh(x);
  S = mk(-f[0] * sh, f[1] * sh);
  R = -f[2] * sh;
}

// sel (optional): per chain 0 = stage list sl, 1 = sl1, -1 = leave the chain unchanged
template <bool ROWS>
__global__ __launch_bounds__(256) void k_apply_stages(const ChainDesc* chains, MatRef Xr, StageList sl0, StageList sl1,
                                                      const int* sel, int slice, const int* slices, int vpb) {
  const ChainDesc& c = chains[blockIdx.y];
  const int n = c.n, N = c.N;
  const int v0 = blockIdx.x * vpb;
  if (v0 >= n) return;
  const int which = sel ? sel[blockIdx.y] : 0;
  if (which < 0) return;
  const StageList& sl = which == 0 ? sl0 : sl1;
  const int nv = min(vpb, n - v0);
  const int l = slices ? slices[blockIdx.y] : slice;
  cplx* X = matptr(c, Xr);
  if (!X) return;

  extern __shared__ double smem[];
  cplx* vec = reinterpret_cast<cplx*>(smem);          // [vpb][n]
  double* evc = smem + 2 * (size_t)vpb * n;           // [N][4] C, S.re, S.im, R

  // plaquette blocks shared by all plaquettes of a (stage, flavour block): one LDS copy
  __shared__ cplx pl_lds[5][4][16];
  const bool uni = c.plaq_uniform != 0;
  const int nplaq_all = N / 4;
  if (uni) {
    for (int e = threadIdx.x; e < min(sl.count, 5) * 64; e += blockDim.x) {
      const int s = e >> 6, b = (e >> 4) & 3, k = e & 15;
      const Stage st = sl.st[s];
      if (st.kind == ST_PLAQ)
        pl_lds[s][b][k] = c.plaq[((((size_t)st.var * 2 + st.group) * 4 + b) * nplaq_all) * 16 + k];
    }
  }
  bool need_ev = false;
  for (int s = 0; s < sl.count; ++s) need_ev |= (sl.st[s].kind == ST_EV);
  if (need_ev) {
    for (int i = threadIdx.x; i < N; i += blockDim.x) {
      double C, R;
      cplx S;
      ev_coeffs(c, l, i, C, S, R);
      evc[4 * i + 0] = C;
      evc[4 * i + 1] = S.re;
      evc[4 * i + 2] = S.im;
      evc[4 * i + 3] = R;
    }
  }
  // load: eight loads per thread in flight before their LDS stores (one memory latency per
  // 2048 entries instead of one per 256)
  {
    const int tot = nv * n;
    for (int i0 = threadIdx.x; i0 < tot; i0 += 8 * blockDim.x) {
      cplx t[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int idx = min(i0 + u * (int)blockDim.x, tot - 1);
        if (!ROWS) {
          const int v = idx / n, k = idx - v * n;
          t[u] = ld(gp(X) + (size_t)(v0 + v) * n + k);
        } else {
          const int k = idx / nv, v = idx - k * nv;
          t[u] = ld(gp(X) + (size_t)k * n + (v0 + v));
        }
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int idx = i0 + u * (int)blockDim.x;
        if (idx < tot) {
          int v, k;
          if (!ROWS) {
            v = idx / n;
            k = idx - v * n;
          } else {
            k = idx / nv;
            v = idx - k * nv;
          }
          vec[v * n + k].re = t[u].re;  // (whole-struct copies from a register array are
          vec[v * n + k].im = t[u].im;  // lowered through a stack slot)
        }
      }
    }
  }
  __syncthreads();

  const int nplaq = N / 4;
  for (int s = 0; s < sl.count; ++s) {
    const Stage st = sl.st[s];
    if (st.kind == ST_MU || st.kind == ST_MUINV) {
      const double* mu = (st.kind == ST_MU) ? c.expmu : c.expmu_inv;
      for (int idx = threadIdx.x; idx < nv * n; idx += blockDim.x) {
        const int v = idx / n, k = idx - v * n;
        vec[v * n + k] = mu[k] * vec[v * n + k];
      }
    } else if (st.kind == ST_EV) {
      if (st.mode == MODE_N) ev_stage<MODE_N>(vec, evc, st.power, n, N, nv);
      else if (st.mode == MODE_T) ev_stage<MODE_T>(vec, evc, st.power, n, N, nv);
      else ev_stage<MODE_C>(vec, evc, st.power, n, N, nv);
    } else if (st.kind == ST_BOND) {
      const cplx* B0 = c.bond_blocks + (size_t)st.var * c.nbonds * 16;
      const int q0 = c.grp_off[st.group], q1 = c.grp_off[st.group + 1];
      if (st.mode == MODE_N) bond_stage<MODE_N>(vec, B0, c.bond_sites, q0, q1, n, N, nv);
      else if (st.mode == MODE_T) bond_stage<MODE_T>(vec, B0, c.bond_sites, q0, q1, n, N, nv);
      else bond_stage<MODE_C>(vec, B0, c.bond_sites, q0, q1, n, N, nv);
    } else {  // ST_PLAQ
      const cplx* P0 = c.plaq + (((size_t)st.var * 2 + st.group) * 4) * nplaq * 16;
      const int* sites = c.plaq_sites + (size_t)st.group * nplaq * 4;
      const cplx* Puni = &pl_lds[s][0][0];
      if (uni) {
        if (st.mode == MODE_N) plaq_stage<MODE_N, true>(vec, P0, Puni, sites, n, N, nv);
        else if (st.mode == MODE_T) plaq_stage<MODE_T, true>(vec, P0, Puni, sites, n, N, nv);
        else plaq_stage<MODE_C, true>(vec, P0, Puni, sites, n, N, nv);
      } else {
        if (st.mode == MODE_N) plaq_stage<MODE_N, false>(vec, P0, Puni, sites, n, N, nv);
        else if (st.mode == MODE_T) plaq_stage<MODE_T, false>(vec, P0, Puni, sites, n, N, nv);
        else plaq_stage<MODE_C, false>(vec, P0, Puni, sites, n, N, nv);
      }
    }
    __syncthreads();
  }
  // store
  if (!ROWS) {
    for (int idx = threadIdx.x; idx < nv * n; idx += blockDim.x) {
      const int v = idx / n, k = idx - v * n;
      st(gp(X) + (size_t)(v0 + v) * n + k, vec[v * n + k]);
    }
  } else {
    for (int idx = threadIdx.x; idx < nv * n; idx += blockDim.x) {
      const int k = idx / nv, v = idx - k * nv;
      st(gp(X) + (size_t)k * n + (v0 + v), vec[v * n + k]);
    }
  }
}

static Stage mkst(int kind, int mode, int var, int group, double power) {
  Stage s;
  s.kind = kind;
  s.mode = mode;
  s.var = var;
  s.group = group;
  s.power = power;
  return s;
}

static hipError_t run_stages(hipStream_t s, const ChainDesc* chains, int nchains, int nmax, int Nmax, MatRef X,
                             bool rows, const StageList& sl, int slice, const int* slices,
                             const StageList* sl1 = nullptr, const int* sel = nullptr) {
  const StageList& s1 = sl1 ? *sl1 : sl;
  // tile of vectors per block: LDS <= 64 KB, and narrower tiles until the launch has a block
  // per CU (a lone chain at n = 400 would otherwise run 100 / 50 blocks); rows stay >= 2 wide
  // (32-byte pieces of each column)
  int vpb = rows ? 8 : 4;
  while (vpb > 1 && (size_t)vpb * nmax * 16 > 65536) vpb >>= 1;
  while (vpb > (rows ? 2 : 1) && (long long)((nmax + vpb - 1) / vpb) * nchains < 256) vpb >>= 1;
  const size_t lds = (size_t)vpb * nmax * 16 + (size_t)Nmax * 32;
  dim3 grid((nmax + vpb - 1) / vpb, nchains);
  {
    hipError_t e = ensure_dynamic_lds(rows ? (const void*)k_apply_stages<true> : (const void*)k_apply_stages<false>,
                                      lds);
    if (e != hipSuccess) return e;
  }
  if (rows)
    hipLaunchKernelGGL(k_apply_stages<true>, grid, dim3(256), lds, s, chains, X, sl, s1, sel, slice, slices, vpb);
  else
    hipLaunchKernelGGL(k_apply_stages<false>, grid, dim3(256), lds, s, chains, X, sl, s1, sel, slice, slices, vpb);
  return hipGetLastError();
}

// Checkerboard factor order (slice_matrices.jl:101-226 Assaad, :233-393 generic). With the
// groups numbered 0 .. G-1 (Assaad: [A, B] = [0, 1]; generic: the greedy bond groups, group 0
// taking the full step):
//   B = h[G-1] ... h[1] H[0] h[1] ... h[G-1] mu eV        (h = half step, H = full step)
// Left application runs the factors right to left on columns, right application left to
// right on rows (transposed mixes); B^-1 and B^dagger follow by inverting / adjoining each.
static hipError_t stage_list(BOp op, int ngroups, StageList& sl, bool& rows) {
  const bool generic = ngroups > 0;
  const int G = generic ? ngroups : 2;
  const int kind = generic ? ST_BOND : ST_PLAQ;
  if (G > kMaxBondGroups) return hipErrorInvalidValue;
  sl.count = 0;
  rows = false;
  auto push = [&](int k, int mode, int var, int group, double power) { sl.st[sl.count++] = mkst(k, mode, var, group, power); };
  // the hopping part h[G-1] ... h[1] H[0] h[1] ... h[G-1] in application order (its two halves
  // are mirror images, so the order is the same from either side)
  auto hops = [&](int mode, int half, int full) {
    for (int g = G - 1; g >= 1; --g) push(kind, mode, half, g, 0);
    push(kind, mode, full, 0, 0);
    for (int g = 1; g < G; ++g) push(kind, mode, half, g, 0);
  };
  switch (op) {
    case B_LEFT:
      push(ST_EV, MODE_N, 0, 0, 1.0);
      push(ST_MU, MODE_N, 0, 0, 0);
      hops(MODE_N, PV_HALF, PV_FULL);
      break;
    case B_INV_LEFT:
      hops(MODE_N, PV_HALF_INV, PV_FULL_INV);
      push(ST_MUINV, MODE_N, 0, 0, 0);
      push(ST_EV, MODE_N, 0, 0, -1.0);
      break;
    case B_DAG_LEFT:
      hops(MODE_C, PV_HALF, PV_FULL);
      push(ST_MU, MODE_N, 0, 0, 0);
      push(ST_EV, MODE_C, 0, 0, 1.0);
      break;
    case B_RIGHT:
      rows = true;
      hops(MODE_T, PV_HALF, PV_FULL);
      push(ST_MU, MODE_N, 0, 0, 0);
      push(ST_EV, MODE_T, 0, 0, 1.0);
      break;
    case B_INV_RIGHT:
      rows = true;
      push(ST_EV, MODE_T, 0, 0, -1.0);
      push(ST_MUINV, MODE_N, 0, 0, 0);
      hops(MODE_T, PV_HALF_INV, PV_FULL_INV);
      break;
    case B_EFF_RIGHT:  // X h[G-1] ... h[0] (fermion_measurements.jl:1132-1135)
      rows = true;
      for (int g = G - 1; g >= 0; --g) push(kind, MODE_T, PV_HALF, g, 0);
      break;
    case B_EFF_INV_LEFT:  // h_inv[0] ... h_inv[G-1] X (fermion_measurements.jl:1136-1139)
      for (int g = G - 1; g >= 0; --g) push(kind, MODE_N, PV_HALF_INV, g, 0);
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipSuccess;
}

hipError_t launch_apply_B_chkr(hipStream_t s, const ChainDesc* chains, int nchains, int nmax, int Nmax, MatRef X,
                               BOp op, int slice, const int* slices, int ngroups) {
  StageList sl;
  bool rows = false;
  hipError_t e = stage_list(op, ngroups, sl, rows);
  if (e != hipSuccess) return e;
  return run_stages(s, chains, nchains, nmax, Nmax, X, rows, sl, slice, slices);
}

hipError_t launch_apply_B_chkr_sel(hipStream_t s, const ChainDesc* chains, int nchains, int nmax, int Nmax, MatRef X,
                                   BOp op0, BOp op1, const int* sel, const int* slices, int ngroups) {
  StageList sl0, sl1;
  bool rows0 = false, rows1 = false;
  hipError_t e = stage_list(op0, ngroups, sl0, rows0);
  if (e == hipSuccess) e = stage_list(op1, ngroups, sl1, rows1);
  if (e != hipSuccess) return e;
  if (rows0 != rows1) return hipErrorInvalidValue;
  return run_stages(s, chains, nchains, nmax, Nmax, X, rows0, sl0, 0, slices, &sl1, sel);
}

// Only the eV factor (used by the dense no-checkerboard path).
hipError_t launch_apply_eV(hipStream_t s, const ChainDesc* chains, int nchains, int nmax, int Nmax, MatRef X,
                           bool rows, int mode, double power, int slice, const int* slices) {
  StageList sl;
  sl.count = 1;
  sl.st[0] = mkst(ST_EV, mode, 0, 0, power);
  return run_stages(s, chains, nchains, nmax, Nmax, X, rows, sl, slice, slices);
}

}  // namespace dqmc
