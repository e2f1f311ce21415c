// Blocked Householder QR on a chain group, and the solve that uses it.
//
// k_qr_blocked factors A P0 = Q R panel by panel (compact WY, Q_p = I - V_p T_p V_p^H,
// LAPACK zgeqrf/zlarft conventions). P0 is either the identity or the pre-pivoting
// order: columns sorted once by decreasing initial norm (ties by index). For the
// column-graded products DQMC stabilises, (B...B U) diag(D), QR with pre-pivoting is
// as stable as column-pivoted QR (Bai, Lee, Li, Xu, "Stable solutions of linear systems
// involving long chain of matrix multiplications", Linear Algebra Appl. 435 (2011)),
// while needing one group barrier per panel instead of one per column.
//
// The output uses the layout of k_qrp (columns never move: logical column k lives in
// physical column qr_perm[k]; reflector k below the diagonal of that column, R on and
// above it; qr_tau[k]; D[k] = |Re R_kk|), so k_form_T / k_form_Q form U and T of
// decompose_udt (linalg.jl:20-39) unchanged. The T factors of all panels are kept in
// qr_Tf for k_qr_solve.
//
// Per panel p, with one-panel look-ahead: workgroup 0 applies Q_p^H to panel p+1's columns
// and factors panel p+1 in registers (zlarfg per column, applied to the rest of the panel,
// with the zlarft inner products formed along the way) while the other workgroups apply
// Q_p^H to the columns beyond (one wave per column); one group barrier per panel.
#include "chain.hpp"

namespace dqmc {

constexpr int kQrMaxN = 1024;
constexpr int kQrMaxNb = 16;

typedef double v4d __attribute__((ext_vector_type(4)));
// one v_mfma_f64_16x16x4: c + A B for this lane's A(i, k) = a, B(k, j) = b
__device__ inline v4d mfma4(double a, double b, v4d c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
// LDS column stride of the panel V and of LDS-resident column blocks: >= n and 1 mod 16, so
// that the 16 columns a quarter-wave reads at one row (128-bit accesses) fall on 16 different
// groups of four banks
inline int mfma_ld(int nmax) { return nmax + ((17 - (nmax & 15)) & 15); }

__device__ inline cplx shfl_c(cplx v, int src) { return mk(__shfl(v.re, src, 64), __shfl(v.im, src, 64)); }
__device__ inline cplx read_lane_c(cplx v, int lane) { return mk(read_lane(v.re, lane), read_lane(v.im, lane)); }

// Panel staging between A (global) and P (LDS, [NB][n], rows k0..n-1 of each column);
// logical column k is physical column perm[k] (perm == nullptr: the identity).
// Each thread moves one row of all NB columns per step: the NB loads are issued before
// any store, so memory latency is paid once per row rather than once per element.
// mode 0: P <- A as is; mode 1: P <- unit lower trapezoid (V with its implicit ones and
// zeros above the diagonal), columns t >= kb zero-filled.
template <int NB>
__device__ inline void stage_panel(cplx* P, const gcplx* A, const int* perm, int k0, int kb, int n, int tid,
                                   int mode, int cs = 0) {
  if (cs == 0) cs = n;  // LDS column stride of P
  const gcplx* src[NB];
#pragma unroll
  for (int t = 0; t < NB; ++t) {
    const int k = k0 + min(t, kb - 1);
    src[t] = A + (size_t)(perm ? perm[k] : k) * n;
  }
  for (int r = k0 + tid; r < n; r += 256) {
    cplx v[NB];
#pragma unroll
    for (int t = 0; t < NB; ++t) v[t] = ld(src[t] + r);
#pragma unroll
    for (int t = 0; t < NB; ++t) {
      cplx val = v[t];
      if (mode == 1) val = (t >= kb || r < k0 + t) ? mk(0, 0) : (r == k0 + t) ? mk(1, 0) : val;
      P[(size_t)t * cs + r] = val;
    }
  }
}

// Workgroup 0's panel work keeps the panel in registers: thread tid owns rows r = tid + 256 q
// (q < RQ) of the NB panel columns. Every cross-thread sum goes through one butterfly
// reduce-scatter per wave and one LDS exchange, so a panel column costs one barrier.
// Selections below are written without data-dependent branches around loads: a guard
// around each LDS read makes the compiler wait for it before the next one.
template <int NB, int RQ>
__global__ __launch_bounds__(256) void k_qr_blocked(const ChainDesc* chains, MatRef Ar, VecRef Dr, int prepivot,
                                                    int ps, const int* bmap) {
  constexpr int V = 2 * NB;               // doubles per butterfly
  constexpr int LS = NB == 8 ? 2 : 1;     // after it, lane l holds value l >> LS
  // grid (W, chains), or with a block map a 1-D grid where block b serves chain bmap[b] & 0xFFFF
  // as workgroup (bmap[b] >> 16) & 0xFF of (bmap[b] >> 24) (workgroup counts per chain)
  const int bm = bmap ? bmap[blockIdx.x] : 0;
  const ChainDesc& c = chains[bmap ? (bm & 0xFFFF) : blockIdx.y];
  const int n = c.n;
  const int W = bmap ? (bm >> 24) & 0xFF : gridDim.x, w = bmap ? (bm >> 16) & 0xFF : blockIdx.x;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, nwaves = blockDim.x >> 6;
  gcplx* A = gp(matptr(c, Ar));
  gdouble* D = gp(vecptr(c, Dr));
  const GroupSync g{c.bar_count, c.bar_gen, c.error, (unsigned int)W};
  // the matrix-core trailing update pays off when each trailing workgroup holds many columns
  // (at least 128 at the start); with few, its tiles are half empty and the block that panel
  // p+2 needs sits on one wave
  const bool mt = NB == 8 && W > 1 && n >= 128 * (W - 1);

  extern __shared__ double smem[];
  cplx* P = reinterpret_cast<cplx*>(smem);  // [NB][ps] panel V (rows k0..n-1 used)
  cplx* wscr = P + (size_t)NB * ps;          // [4][8][16] per-wave W of the matrix-core trailing update
  __shared__ int perm[kQrMaxN], lpos[kQrMaxN];
  __shared__ cplx Tl[NB * NB];  // Tl[j * NB + t] = T(j, t)
  __shared__ cplx taul[NB];
  __shared__ double betal[NB];
  __shared__ double red[2][4][V];     // per-wave sums of a panel step, by step parity
  __shared__ cplx piv[2][NB];         // the pivot row of a panel step, by step parity
  __shared__ double redW[4][NB][V + 1];  // per-wave V^H a of NB columns (padded: no bank conflicts)

  // ---- column order
  if (prepivot) {
    for (int col = w + W * wave; col < n; col += W * nwaves) {
      const gcplx* a = A + (size_t)col * n;
      double acc = 0.0;
      for (int k = lane; k < n; k += 64) acc += abs2(ld(a + k));
      acc = wave_sum(acc);
      if (lane == 0) gp(c.qr_vn)[col] = sqrt(acc);
    }
    group_sync(g);
    if (group_failed(g)) return;
    double* vn = reinterpret_cast<double*>(P);  // n doubles, before the panel buffer is used
    for (int k = tid; k < n; k += blockDim.x) vn[k] = gp(c.qr_vn)[k];
    __syncthreads();
    for (int col = tid; col < n; col += blockDim.x) {
      const double v = vn[col];
      int r = 0;
      for (int q = 0; q < n; ++q) {
        const double u = vn[q];
        r += (u > v || (u == v && q < col)) ? 1 : 0;
      }
      lpos[col] = r;
    }
    __syncthreads();
    for (int col = tid; col < n; col += blockDim.x) perm[lpos[col]] = col;
  } else {
    for (int k = tid; k < n; k += blockDim.x) perm[k] = lpos[k] = k;
  }
  __syncthreads();

  cplx x[RQ][NB];  // this thread's rows of the panel being factored

  // rows >= kr of logical columns kc .. kc+kbc-1 into x (everything else zero)
  auto load_regs = [&](int kc, int kbc, int kr) __attribute__((always_inline)) {
#pragma unroll
    for (int t = 0; t < NB; ++t) {
      const gcplx* a = A + (size_t)perm[kc + min(t, kbc - 1)] * n;
#pragma unroll
      for (int q = 0; q < RQ; ++q) {
        const int r = tid + 256 * q;
        const gcplx* src = a + min(r, n - 1);
        const cplx v = mk(src->re, src->im);
        x[q][t] = (t < kbc && r >= kr && r < n) ? v : mk(0, 0);
      }
    }
  };

  // ---- factor the panel at logical column k0 (panel index p) held in x; rows >= kw of x are
  // valid and written back. Per column t (zlarfg, then H^H on the columns after it):
  //   one butterfly sums |x_t|^2 and x_t^H a_u (u > t) over the rows below the pivot, plus
  //   v_j^H x_t (j < t) for zlarft; the pivot row goes through LDS; after one barrier every
  //   thread forms beta, tau and 1/(alpha - beta) from the same sums and updates its rows:
  //   v^H a_u = a_u(i) + conj(1/(alpha - beta)) x_t^H a_u, and wave 0 stores
  //   S(j, t) = v_j^H v_t = conj(v_j(i)) + (1/(alpha - beta)) v_j^H x_t.
  // Every call site of this and of the block update below is single, and the kernel's code
  // fits the instruction cache: the panel loop runs without instruction fetch stalls.
  auto factor_regs = [&](int k0, int p, int kw) __attribute__((always_inline)) {
    const int kb = min(NB, n - k0);
    gcplx* Tg = gp(c.qr_Tf) + (size_t)p * NB * NB;
    static_for<0, NB>([&](auto tc) __attribute__((always_inline)) {
      constexpr int t = decltype(tc)::value, par = t & 1;
      const int i = k0 + t;
      double v[V];
#pragma unroll
      for (int e = 0; e < V; ++e) v[e] = 0.0;
#pragma unroll
      for (int q = 0; q < RQ; ++q) {
        const int r = tid + 256 * q;
        const cplx xt = (r > i) ? x[q][t] : mk(0, 0);
        v[0] += abs2(xt);
#pragma unroll
        for (int u = t + 1; u < NB; ++u) {
          const int e = 1 + 2 * (u - t - 1);
          v[e] = fma(xt.re, x[q][u].re, fma(xt.im, x[q][u].im, v[e]));
          v[e + 1] = fma(xt.re, x[q][u].im, fma(-xt.im, x[q][u].re, v[e + 1]));
        }
#pragma unroll
        for (int j = 0; j < t; ++j) {  // v_j^H x_t below the pivot, for S(j, t)
          const int e = 1 + 2 * (NB - 1 - t) + 2 * j;
          v[e] = fma(x[q][j].re, xt.re, fma(x[q][j].im, xt.im, v[e]));
          v[e + 1] = fma(x[q][j].re, xt.im, fma(-x[q][j].im, xt.re, v[e + 1]));
        }
      }
#pragma unroll
      for (int q = 0; q < RQ; ++q)
        if (tid + 256 * q == i) {  // the owner of the pivot row (all columns: v_j(i) for j < t too)
#pragma unroll
          for (int u = 0; u < NB; ++u) {
            piv[par][u].re = x[q][u].re;
            piv[par][u].im = x[q][u].im;
          }
        }
      wave_reduce_scatter<V>(v);
      if ((lane & ((1 << LS) - 1)) == 0) red[par][wave][lane >> LS] = v[0];
      __syncthreads();
      const double xsq = (red[par][0][0] + red[par][1][0]) + (red[par][2][0] + red[par][3][0]);
      const cplx alpha = mk(piv[par][t].re, piv[par][t].im);
      cplx tau, sc;
      double beta;
      // zlarfg: H^H (alpha, x) = (beta, 0), v = (1, x / (alpha - beta))
      if (xsq == 0.0 && alpha.im == 0.0) {
        tau = mk(0, 0);
        beta = alpha.re;
        sc = mk(1, 0);
      } else {
        beta = -copysign(sqrt(alpha.re * alpha.re + alpha.im * alpha.im + xsq), alpha.re);
        const double ib = 1.0 / beta;  // two divisions on the per-column critical path, not five
        tau = mk((beta - alpha.re) * ib, -alpha.im * ib);
        const cplx d = alpha - mk(beta, 0);  // |d|^2 <= 4 (|alpha|^2 + xsq): no new overflow
        const double id = 1.0 / (d.re * d.re + d.im * d.im);
        sc = mk(d.re * id, -d.im * id);
      }
      // lane k forms f for column t+1+k (its sums read by lane), then every lane takes them
      // from scalar registers
      cplx fl;
      {
        const int k = min(lane, NB - 1 - t > 0 ? NB - 2 - t : 0), e = 1 + 2 * k;
        const cplx raw = mk((red[par][0][e] + red[par][1][e]) + (red[par][2][e] + red[par][3][e]),
                            (red[par][0][e + 1] + red[par][1][e + 1]) + (red[par][2][e + 1] + red[par][3][e + 1]));
        const int u = min(t + 1 + k, NB - 1);
        cplx dot = mk(piv[par][u].re, piv[par][u].im);
        cfma_cj(dot, sc, raw);
        fl = conj(tau) * dot;
      }
      cplx f[NB];
#pragma unroll
      for (int u = t + 1; u < NB; ++u) f[u] = read_lane_c(fl, u - t - 1);
#pragma unroll
      for (int q = 0; q < RQ; ++q) {
        const int r = tid + 256 * q;
        const cplx xt = x[q][t];
        const cplx vt = (r > i) ? sc * xt : (r == i) ? mk(1, 0) : mk(0, 0);
#pragma unroll
        for (int u = t + 1; u < NB; ++u) x[q][u] = x[q][u] - f[u] * vt;
        x[q][t] = (r > i) ? vt : (r == i) ? mk(beta, 0) : xt;
      }
      if (wave == 0 && lane < t) {  // lane j: S(j, t) = v_j^H v_t = conj(v_j(i)) + sc v_j^H x_t
        const int e = 1 + 2 * (NB - 1 - t) + 2 * lane;
        const cplx raw = mk((red[par][0][e] + red[par][1][e]) + (red[par][2][e] + red[par][3][e]),
                            (red[par][0][e + 1] + red[par][1][e + 1]) + (red[par][2][e + 1] + red[par][3][e + 1]));
        cplx sj = conj(mk(piv[par][lane].re, piv[par][lane].im));
        cfma(sj, sc, raw);
        Tl[lane * NB + t] = sj;
      }
      if (tid == 0) {
        taul[t] = tau;
        betal[t] = beta;
      }
    });
    if (tid < kb) {
      D[k0 + tid] = fabs(betal[tid]);
      st(gp(c.qr_tau) + k0 + tid, taul[tid]);
    }
    // ---- zlarft (forward, columnwise) from S: T(t,t) = tau_t, T(0:t, t) = -tau_t T(0:t, 0:t) S(0:t, t);
    // wave 0, lane j keeps row j of T in registers (T(j, l) = 0 for j > l)
    if (wave == 0) {
      cplx trow[NB];
#pragma unroll
      for (int l = 0; l < NB; ++l) trow[l] = mk(0, 0);
      // no branch on t < kb around the LDS reads (that would wait for each read in turn):
      // the S entries and tau of the zero columns past kb are zero, and a select keeps them 0
#pragma unroll
      for (int t = 0; t < NB; ++t) {
        cplx acc = mk(0, 0);
#pragma unroll
        for (int l = 0; l < t; ++l) cfma(acc, trow[l], Tl[l * NB + t]);
        const cplx tt = taul[t];
        const cplx v = (lane < t) ? -(tt * acc) : (lane == t) ? tt : mk(0, 0);
        trow[t] = (t < kb) ? v : mk(0, 0);
      }
      if (lane < NB) {
#pragma unroll
        for (int t = 0; t < NB; ++t) {
          Tl[lane * NB + t] = trow[t];
          st(Tg + lane * NB + t, trow[t]);
        }
      }
    }
    // x back to A (rows >= kw: R above and on the diagonal, reflectors below) and, as the
    // unit lower trapezoid V, to P for the next panel's update
#pragma unroll
    for (int t = 0; t < NB; ++t) {
      gcplx* a = A + (size_t)perm[k0 + min(t, kb - 1)] * n;
#pragma unroll
      for (int q = 0; q < RQ; ++q) {
        const int r = tid + 256 * q;
        const cplx xv = x[q][t];
        if (r >= k0 && r < n) {
          const cplx pv = (t >= kb || r < k0 + t) ? mk(0, 0) : (r == k0 + t) ? mk(1, 0) : xv;
          P[(size_t)t * ps + r].re = pv.re;
          P[(size_t)t * ps + r].im = pv.im;
        }
        if (t < kb && r >= kw && r < n) {
          a[r].re = xv.re;
          a[r].im = xv.im;
        }
      }
    }  // the group barrier that follows every call publishes P, Tl and A
  };

  // ---- x <- (I - V T^H V^H) x for NB columns held in x (rows >= k0), with V (rows >= k0) in P,
  // this thread's rows of it in pj, and T in Tl: W = V^H x by one butterfly per column, then
  // every wave forms Z = T^H W itself (entry u NB + j on lane (u NB + j) mod 64) and the row
  // updates take Z from scalar registers. V is read from LDS once per panel, not per column.
  cplx pj[RQ][NB];
  auto load_pj = [&](int k0) __attribute__((always_inline)) {
#pragma unroll
    for (int q = 0; q < RQ; ++q) {
      const int r = tid + 256 * q;
#pragma unroll
      for (int j = 0; j < NB; ++j) {
        const cplx* src = P + (size_t)j * ps + min(r, n - 1);
        const cplx pv = mk(src->re, src->im);
        pj[q][j] = (r >= k0 && r < n) ? pv : mk(0, 0);
      }
    }
  };
  auto apply_regs = [&]() __attribute__((always_inline)) {
    constexpr int ZE = NB * NB / 64;  // Z entries per lane
#pragma unroll
    for (int u = 0; u < NB; ++u) {
      double v[V];
#pragma unroll
      for (int e = 0; e < V; ++e) v[e] = 0.0;
#pragma unroll
      for (int q = 0; q < RQ; ++q)
#pragma unroll
        for (int j = 0; j < NB; ++j) {
          v[2 * j] = fma(pj[q][j].re, x[q][u].re, fma(pj[q][j].im, x[q][u].im, v[2 * j]));
          v[2 * j + 1] = fma(pj[q][j].re, x[q][u].im, fma(-pj[q][j].im, x[q][u].re, v[2 * j + 1]));
        }
      wave_reduce_scatter<V>(v);
      if ((lane & ((1 << LS) - 1)) == 0) redW[wave][u][lane >> LS] = v[0];
    }
    __syncthreads();
    cplx zr[ZE];
#pragma unroll
    for (int m = 0; m < ZE; ++m) {
      const int idx = lane + 64 * m, u = idx / NB, j = idx % NB;
      cplx z = mk(0, 0);
#pragma unroll
      for (int l = 0; l < NB; ++l) {
        const cplx wl = mk((redW[0][u][2 * l] + redW[1][u][2 * l]) + (redW[2][u][2 * l] + redW[3][u][2 * l]),
                           (redW[0][u][2 * l + 1] + redW[1][u][2 * l + 1]) +
                               (redW[2][u][2 * l + 1] + redW[3][u][2 * l + 1]));
        cfma_cj(z, Tl[l * NB + j], wl);
      }
      zr[m] = z;
    }
#pragma unroll
    for (int u = 0; u < NB; ++u)
#pragma unroll
      for (int j = 0; j < NB; ++j) {
        const int idx = u * NB + j;
        const cplx z = read_lane_c(zr[idx / 64], idx % 64);
#pragma unroll
        for (int q = 0; q < RQ; ++q) {
          cplx a = x[q][u];
          a.re = fma(-pj[q][j].re, z.re, fma(pj[q][j].im, z.im, a.re));
          a.im = fma(-pj[q][j].re, z.im, fma(-pj[q][j].im, z.re, a.im));
          x[q][u] = a;
        }
      }
  };
  // rows >= kr of x back to logical columns kc .. kc+kbc-1
  auto store_regs = [&](int kc, int kbc, int kr) __attribute__((always_inline)) {
#pragma unroll
    for (int t = 0; t < NB; ++t) {
      gcplx* a = A + (size_t)perm[kc + min(t, kbc - 1)] * n;
#pragma unroll
      for (int q = 0; q < RQ; ++q) {
        const int r = tid + 256 * q;
        if (t < kbc && r >= kr && r < n) {
          a[r].re = x[q][t].re;
          a[r].im = x[q][t].im;
        }
      }
    }
  };
  const int ws = W == 1 ? 1 : W - 1, w0 = W == 1 ? 0 : 1;  // trailing workgroups, first one
  // Trailing update of panel p (rows >= k0) on the matrix cores (mt), for workgroups w >= 1: this workgroup's blocks (block b of the columns beyond k1 + NB goes to
  // workgroup 1 + b mod (W-1)) in tiles of two blocks, 16 columns, tile q to wave q mod 4.
  // Per tile: W = V^H X (K = rows, A(i, k) = conj V(r, i), B(k, j) = X(r, col_j)), then
  // Z = T^H W with the vector units (lane (j, k) forms Z(k, j) and Z(k + 4, j), exactly its B
  // operands of the second product), then X -= V Z in 16-row tiles. X streams from global
  // memory twice (loads four steps ahead); V comes from P (stride ps: conflict-free).
  auto trailing_mfma = [&](int k0, int k1) __attribute__((always_inline)) {
    const int li = lane & 15, lk = lane >> 4, vi = min(li, NB - 1);
    const int nbt = (n - k1 - NB + NB - 1) / NB;  // blocks beyond panel p+1
    const int nbw = nbt > w - 1 ? (nbt - (w - 1) + ws - 1) / ws : 0;
    const int nst = (n - k0) >> 2;
    cplx* wsc = wscr + wave * 128;
    for (int q = wave; 2 * q < nbw; q += nwaves) {
      const int j = 2 * q + (li >> 3);
      const int cl = k1 + NB + ((w - 1) + j * ws) * NB + (li & 7);
      const bool valid = j < nbw && cl < n;
      gcplx* a = A + (size_t)perm[valid ? cl : min(k1 + NB, n - 1)] * n;
      v4d wr = {0.0, 0.0, 0.0, 0.0}, wi = {0.0, 0.0, 0.0, 0.0};
      constexpr int PF = 8;  // k-steps of X loaded ahead of the products
      cplx xb[PF];
#pragma unroll
      for (int u = 0; u < PF; ++u) xb[u] = ld(a + k0 + 4 * min(u, nst - 1) + lk);
      for (int s0 = 0; s0 < nst; s0 += PF) {
        cplx xc[PF];
#pragma unroll
        for (int u = 0; u < PF; ++u) xc[u] = xb[u];
#pragma unroll
        for (int u = 0; u < PF; ++u) xb[u] = ld(a + k0 + 4 * min(s0 + PF + u, nst - 1) + lk);
#pragma unroll
        for (int u = 0; u < PF; ++u) {
          if (s0 + u < nst) {
            const cplx v = P[(size_t)vi * ps + k0 + 4 * (s0 + u) + lk];
            const double vr = li < NB ? v.re : 0.0, vm = li < NB ? v.im : 0.0;
            const double xr = valid ? xc[u].re : 0.0, xm = valid ? xc[u].im : 0.0;
            // W += V^H X: re += Vr Xr + Vi Xi, im += Vr Xi - Vi Xr
            wr = mfma4(vr, xr, wr);
            wi = mfma4(vr, xm, wi);
            wr = mfma4(vm, xm, wr);
            wi = mfma4(-vm, xr, wi);
          }
        }
      }
      // lane (li, lk) holds W(lk, li) and W(lk + 4, li); every lane needs W(:, li)
      wsc[lk * 16 + li] = mk(wr[0], wi[0]);
      wsc[(lk + 4) * 16 + li] = mk(wr[1], wi[1]);
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      cplx z0 = mk(0, 0), z1 = mk(0, 0);
#pragma unroll
      for (int l = 0; l < NB; ++l) {
        const cplx wl = wsc[l * 16 + li];
        cfma_cj(z0, Tl[l * NB + lk], wl);
        cfma_cj(z1, Tl[l * NB + lk + 4], wl);
      }
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");  // W read before the next tile writes it
      // X -= V Z over 16-row tiles: element v of lane (li, lk) is row r0 + lk + 4 v, column li
      cplx cb[4], cn[4];  // the next two row tiles' X, loaded ahead
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        cb[v] = ld(a + min(k0 + lk + 4 * v, n - 1));
        cn[v] = ld(a + min(k0 + 16 + lk + 4 * v, n - 1));
      }
      for (int r0 = k0; r0 < n; r0 += 16) {
        v4d dr, di;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          dr[v] = cb[v].re;
          di[v] = cb[v].im;
          cb[v] = cn[v];
        }
#pragma unroll
        for (int v = 0; v < 4; ++v) cn[v] = ld(a + min(r0 + 32 + lk + 4 * v, n - 1));
#pragma unroll
        for (int ks = 0; ks < 2; ++ks) {
          const int ra = r0 + li;
          const cplx a0 = P[(size_t)(4 * ks + lk) * ps + min(ra, n - 1)];
          const double ar = ra < n ? a0.re : 0.0, am = ra < n ? a0.im : 0.0;
          const cplx z = ks ? z1 : z0;
          // X -= V Z: re += -Vr Zr + Vi Zi, im += -Vr Zi - Vi Zr
          dr = mfma4(-ar, z.re, dr);
          di = mfma4(-ar, z.im, di);
          dr = mfma4(am, z.im, dr);
          di = mfma4(-am, z.re, di);
        }
        if (valid) {
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const int row = r0 + lk + 4 * v;
            if (row < n) {
              a[row].re = dr[v];
              a[row].im = di[v];
            }
          }
        }
      }
    }
  };
  // One-panel look-ahead: while every other workgroup applies panel p to the columns beyond
  // panel p+1 (in blocks of NB logical columns, block b to workgroup 1 + b mod (W-1)),
  // workgroup 0 applies it to panel p+1's columns and factors panel p+1, so each panel costs
  // one group barrier and its factorisation overlaps the trailing update. Workgroup 0 keeps
  // V and T of the panel it factored in P and Tl; the others stage them. Iteration p = -1
  // factors panel 0 alone. Each code path has one call site, for the instruction cache.
  for (int p = -1, k1 = 0; k1 < n; ++p, k1 += NB) {  // apply panel p, factor panel p+1 at k1
    const int k0 = k1 - NB;
    if (p >= 0) {
      if (w != 0) {
        stage_panel<NB>(P, A, perm, k0, NB, n, tid, 1, ps);
        for (int e = tid; e < NB * NB; e += blockDim.x) Tl[e] = ld(gp(c.qr_Tf) + (size_t)p * NB * NB + e);
        __syncthreads();
      }
      if (w == 0 || NB != 8 || !mt) load_pj(k0);
    }
    // this workgroup's blocks: trailing ones first (kc >= k1 + NB), then, on workgroup 0,
    // panel p+1 itself, which stays in x for the factorisation
    if (p < 0) {
      if (w == 0) load_regs(0, min(NB, n), 0);
    } else if (NB == 8 && w != 0 && mt) {
      trailing_mfma(k0, k1);
    } else {
      for (int kc = (w >= w0) ? k1 + NB + (w - w0) * NB : n;; kc += ws * NB) {
        const bool next = kc >= n;
        if (next && w != 0) break;
        const int kcc = next ? k1 : kc, kbc = min(NB, n - kcc);
        load_regs(kcc, kbc, k0);
        apply_regs();
        if (next) break;
        store_regs(kcc, kbc, k0);
        __syncthreads();  // redW is reused by the next block
      }
    }
    if (w == 0) factor_regs(k1, p + 1, p >= 0 ? k0 : 0);
    group_sync(g);
    if (group_failed(g)) return;
  }
  if (w == 0) {
    for (int k = tid; k < n; k += blockDim.x) {
      gp(c.qr_perm)[k] = perm[k];
      gp(c.qr_lpos)[k] = lpos[k];
    }
  }
}

// Column layout of k_qr_solve / k_form_Q_blocked: a workgroup owns CB columns in LDS. With
// CB <= 4 the four waves split as CB columns x (4 / CB) row slices; with CB > 4 each wave owns
// CB / 4 columns and all rows, so every V entry read from LDS serves CB / 4 columns.
template <int CB>
struct ColLayout {
  static constexpr int RS = CB >= 4 ? 1 : 4 / CB;   // row slices per column
  static constexpr int CPW = CB >= 4 ? CB / 4 : 1;  // columns per wave
};

// x <- (I - V op(T) V^H) x for this wave's CPW columns xc + (jl0 + c) n (rows >= k0); V [NB][n]
// and T (Tl[j * NB + l] = T(j, l)) in LDS; op = T^H (ADJ, applying Q^H) or T (applying Q).
// part: LDS for the row-slice sums when RS > 1.
template <int NB, int CB, bool ADJ>
__device__ inline void apply_panel_cols(cplx* xc, const cplx* V, const cplx* Tl, cplx* part, int n, int k0, int jl0,
                                        int rs, int ncol, int lane, int wave) {
  using Lay = ColLayout<CB>;
  constexpr int RS = Lay::RS, CPW = Lay::CPW;
  cplx wv[CPW][NB];
#pragma unroll
  for (int c = 0; c < CPW; ++c)
#pragma unroll
    for (int j = 0; j < NB; ++j) wv[c][j] = mk(0, 0);
  for (int k = k0 + lane + 64 * rs; k < n; k += 64 * RS) {
    cplx pj[NB];
#pragma unroll
    for (int j = 0; j < NB; ++j) pj[j] = V[(size_t)j * n + k];
#pragma unroll
    for (int c = 0; c < CPW; ++c) {
      const cplx ak = xc[(size_t)(jl0 + c) * n + k];
#pragma unroll
      for (int j = 0; j < NB; ++j) cfma_cj(wv[c][j], pj[j], ak);
    }
  }
#pragma unroll
  for (int c = 0; c < CPW; ++c) wave_sum_many<NB>(wv[c]);
  if (RS > 1) {
    if (lane == 0) {
#pragma unroll
      for (int j = 0; j < NB; ++j) part[wave * NB + j] = wv[0][j];
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < NB; ++j) {
      cplx sj = mk(0, 0);
#pragma unroll
      for (int q = 0; q < RS; ++q) sj += part[(jl0 + CB * q) * NB + j];
      wv[0][j] = sj;
    }
  }
  // z = op(T) w: lane j < NB forms z_j of every column, then every lane reads them
  const int jz = lane & (NB - 1);
  cplx z[CPW][NB];
#pragma unroll
  for (int c = 0; c < CPW; ++c) {
    cplx zl = mk(0, 0);
#pragma unroll
    for (int l = 0; l < NB; ++l) {
      if (ADJ)
        cfma_cj(zl, Tl[l * NB + jz], wv[c][l]);
      else
        cfma(zl, Tl[jz * NB + l], wv[c][l]);
    }
#pragma unroll
    for (int j = 0; j < NB; ++j) z[c][j] = read_lane_c(zl, j);
  }
  if (jl0 < ncol) {  // wave-uniform; columns past ncol of the group are scratch
    for (int k = k0 + lane + 64 * rs; k < n; k += 64 * RS) {
      cplx pj[NB];
#pragma unroll
      for (int j = 0; j < NB; ++j) pj[j] = V[(size_t)j * n + k];
#pragma unroll
      for (int c = 0; c < CPW; ++c) {
        cplx ak = xc[(size_t)(jl0 + c) * n + k];
#pragma unroll
        for (int j = 0; j < NB; ++j) ak = ak - pj[j] * z[c][j];
        xc[(size_t)(jl0 + c) * n + k] = ak;
      }
    }
  }
}

// B <- R^-1 Q^H B for a matrix factored by k_qr_blocked without pre-pivoting (Ar holds R and
// the reflectors, qr_tau / qr_Tf the rest). Every right-hand side is independent: each
// workgroup owns CB columns of B in LDS, applies the panels' Q_p^H in order, then
// back-substitutes with R; no synchronisation between workgroups. The four waves split
// as CB columns x (4 / CB) row slices; with more than one slice per column, the panel's
// inner products are summed over the slices through LDS.
template <int NB, int CB>
__global__ __launch_bounds__(256) void k_qr_solve(const ChainDesc* chains, MatRef Ar, MatRef Br) {
  using Lay = ColLayout<CB>;
  int bx, by;
  xcd_block(bx, by);
  const ChainDesc& c = chains[by];
  const int n = c.n;
  const int j0 = bx * CB;
  if (j0 >= n) return;
  const int ncol = min(CB, n - j0);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int jl0 = CB >= 4 ? wave * Lay::CPW : wave % CB, rs = CB >= 4 ? 0 : wave / CB;
  const gcplx* A = gp(matptr(c, Ar));
  gcplx* B = gp(matptr(c, Br));
  extern __shared__ double smem[];
  cplx* bc = reinterpret_cast<cplx*>(smem);  // [CB][n]
  cplx* V = bc + (size_t)CB * n;              // [NB][n]
  __shared__ cplx Tl[NB * NB];
  __shared__ cplx Rd[16 * 16];
  __shared__ cplx part[Lay::RS > 1 ? 4 * NB : 1];
  for (int r = tid; r < n; r += 256) {
    cplx v[CB];
#pragma unroll
    for (int q = 0; q < CB; ++q) {
      const gcplx* src = B + (size_t)(j0 + min(q, ncol - 1)) * n + r;
      v[q] = mk(src->re, src->im);  // (a whole-struct copy here is lowered through a stack slot)
    }
#pragma unroll
    for (int q = 0; q < CB; ++q) bc[(size_t)q * n + r] = v[q];
  }
  for (int k0 = 0, p = 0; k0 < n; k0 += NB, ++p) {
    const int kb = min(NB, n - k0);
    __syncthreads();
    stage_panel<NB>(V, A, nullptr, k0, kb, n, tid, 1);  // columns >= kb zero-filled
    for (int e = tid; e < NB * NB; e += blockDim.x) Tl[e] = ld(gp(c.qr_Tf) + (size_t)p * NB * NB + e);
    __syncthreads();
    apply_panel_cols<NB, CB, true>(bc, V, Tl, part, n, k0, jl0, rs, ncol, lane, wave);
  }
  // back substitution R y = z in blocks of 16 rows from the bottom: wave q < CB solves the
  // 16 x 16 diagonal block for column q (lane r holds row lo + r), then every thread takes
  // rows r < lo and subtracts the block's contribution for all CB columns, reading R's
  // block column once (sixteen loads in flight per row)
  for (int hi = n; hi > 0; hi -= 16) {
    const int lo = max(0, hi - 16), bs = hi - lo;
    __syncthreads();
    {
      const int kk = tid >> 4, rr = tid & 15;
      Rd[tid] = (kk < bs && rr < bs) ? ld(A + (size_t)(lo + kk) * n + lo + rr) : mk(0, 0);  // Rd[kk*16+rr] = R(lo+rr, lo+kk)
    }
    __syncthreads();
    for (int q = wave; q < ncol; q += 4) {  // wave-uniform: the 16 x 16 diagonal solve of column q
      cplx* y = bc + (size_t)q * n;
      const int r = lo + (lane & 15);
      cplx zr = (lane < bs) ? y[r] : mk(0, 0);
      for (int k = bs - 1; k >= 0; --k) {
        const cplx yk = shfl_c(zr, k) * cinv(Rd[k * 16 + k]);
        if (lane < k) zr = zr - Rd[k * 16 + (lane & 15)] * yk;
        if (lane == k) zr = yk;
      }
      if (lane < bs) y[r] = zr;
    }
    __syncthreads();
    for (int r = tid; r < lo; r += 256) {
      cplx rv[16];
#pragma unroll
      for (int k = 0; k < 16; ++k) rv[k] = ld(A + (size_t)(lo + min(k, bs - 1)) * n + r);
#pragma unroll
      for (int q = 0; q < CB; ++q) {
        cplx* y = bc + (size_t)q * n;
        cplx acc = y[r];
#pragma unroll
        for (int k = 0; k < 16; ++k) {
          const cplx yk = y[min(lo + k, n - 1)];
          acc = acc - rv[k] * ((k < bs) ? yk : mk(0, 0));
        }
        y[r] = acc;
      }
    }
  }
  __syncthreads();
  for (int r = tid; r < n; r += 256) {
    cplx v[CB];
#pragma unroll
    for (int q = 0; q < CB; ++q) v[q] = bc[(size_t)q * n + r];
#pragma unroll
    for (int q = 0; q < CB; ++q)
      if (q < ncol) {
        gcplx* d = B + (size_t)(j0 + q) * n + r;
        d->re = v[q].re;
        d->im = v[q].im;
      }
  }
}

// Q = Q_1 Q_2 ... Q_P of a factorisation by k_qr_blocked, Q_p = I - V_p T_p V_p^H
// (zungqr's result, formed from the compact-WY panels instead of one reflector at a time).
// Each workgroup owns CB columns of Q in LDS, starting from the identity, and applies the
// panels from the last one that reaches its columns down to the first; the waves split as
// in k_qr_solve. Reflector k lives in physical column qr_perm[k] below its diagonal.
template <int NB, int CB>
__global__ __launch_bounds__(256) void k_form_Q_blocked(const ChainDesc* chains, MatRef Ar, MatRef Qr) {
  using Lay = ColLayout<CB>;
  int bx, by;
  xcd_block(bx, by);
  const ChainDesc& c = chains[by];
  const int n = c.n;
  const int j0 = bx * CB;
  if (j0 >= n) return;
  const int ncol = min(CB, n - j0);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int jl0 = CB >= 4 ? wave * Lay::CPW : wave % CB, rs = CB >= 4 ? 0 : wave / CB;
  const gcplx* A = gp(matptr(c, Ar));
  gcplx* Q = gp(matptr(c, Qr));
  extern __shared__ double smem[];
  cplx* xc = reinterpret_cast<cplx*>(smem);  // [CB][n]
  cplx* V = xc + (size_t)CB * n;              // [NB][n]
  __shared__ cplx Tl[NB * NB];
  __shared__ cplx part[Lay::RS > 1 ? 4 * NB : 1];
  for (int r = tid; r < n; r += 256)
#pragma unroll
    for (int q = 0; q < CB; ++q) {
      xc[(size_t)q * n + r].re = (r == j0 + q) ? 1.0 : 0.0;
      xc[(size_t)q * n + r].im = 0.0;
    }
  const int jmax = j0 + ncol - 1;  // panels starting beyond it leave these columns alone
  for (int p = jmax / NB; p >= 0; --p) {
    const int k0 = p * NB, kb = min(NB, n - k0);
    __syncthreads();
    stage_panel<NB>(V, A, c.qr_perm, k0, kb, n, tid, 1);  // columns >= kb zero-filled
    for (int e = tid; e < NB * NB; e += blockDim.x) Tl[e] = ld(gp(c.qr_Tf) + (size_t)p * NB * NB + e);
    __syncthreads();
    apply_panel_cols<NB, CB, false>(xc, V, Tl, part, n, k0, jl0, rs, ncol, lane, wave);
  }
  __syncthreads();
  for (int r = tid; r < n; r += 256) {
    cplx v[CB];
#pragma unroll
    for (int q = 0; q < CB; ++q) v[q] = xc[(size_t)q * n + r];
#pragma unroll
    for (int q = 0; q < CB; ++q)
      if (q < ncol) {
        gcplx* d = Q + (size_t)(j0 + q) * n + r;
        d->re = v[q].re;
        d->im = v[q].im;
      }
  }
}

// ---------------------------------------------------------------------------
// Panel application on the FP64 matrix cores for LDS-resident column blocks (k_qr_solve_mfma,
// k_form_Q_mfma). The workgroup owns CB <= 16 columns of X (one 16-column MFMA tile) and
// applies one 8-column compact-WY panel, X <- (I - V op(T) V^H) X on the rows >= k0:
//   W = V^H X (8 x CB): v_mfma_f64_16x16x4 with A(i, k) = conj V(r, i) (rows 8..15 of the
//     tile zero) and B(k, j) = X(r, j), the rows split into four contiguous ranges, one per
//     wave; waves 1-3 leave their partial W in LDS, wave 0 adds them, forms Z = op(T) W with
//     the vector units and leaves Z in LDS;
//   X -= V Z: 16-row tiles of X spread over the waves, A(i, k) = V(r0 + i, k), B(k, j) =
//     Z(k, j), two k-steps.
// A complex product is four real MFMAs. Lane layout of v_mfma_f64_16x16x4: A(i, k) and
// B(k, j) on lane i + 16 k (resp. j + 16 k); D(i, j) element v on lane j + 16 (i - 4 v).
// LDS: X [CB][ld], V [8][ld] (unit lower trapezoid from stage_panel), 3 x 8 x 16 scratch.
// ---------------------------------------------------------------------------
constexpr int kMfmaScratch = 3 * 8 * 16;  // complex entries of the panel scratch

template <int CB, bool ADJ>
__device__ inline void apply_panel_mfma(cplx* xc, const cplx* V, const cplx* Tl, cplx* scr, int n, int cs,
                                        int k0) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int li = lane & 15, lk = lane >> 4;
  const int vi = min(li, 7), xj = min(li, CB - 1);
  const int nst = (n - k0) >> 2;  // n - k0 is a multiple of 4
  const int sa = (nst * wave) >> 2, sb = (nst * (wave + 1)) >> 2;
  v4d wr = {0.0, 0.0, 0.0, 0.0}, wi = {0.0, 0.0, 0.0, 0.0};
  for (int st = sa; st < sb; ++st) {
    const int r = k0 + 4 * st + lk;
    const cplx v = V[(size_t)vi * cs + r];
    const cplx x = xc[(size_t)xj * cs + r];
    const double vr = li < 8 ? v.re : 0.0, vm = li < 8 ? v.im : 0.0;
    const double xr = li < CB ? x.re : 0.0, xm = li < CB ? x.im : 0.0;
    // W += V^H X: re += Vr Xr + Vi Xi, im += Vr Xi - Vi Xr
    wr = mfma4(vr, xr, wr);
    wi = mfma4(vr, xm, wi);
    wr = mfma4(vm, xm, wr);
    wi = mfma4(-vm, xr, wi);
  }
  // lane (li, lk) holds W(lk, li) (element 0) and W(lk + 4, li) (element 1)
  if (wave > 0) {
    cplx* q = scr + (wave - 1) * 128;
    q[lk * 16 + li] = mk(wr[0], wi[0]);
    q[(lk + 4) * 16 + li] = mk(wr[1], wi[1]);
  }
  __syncthreads();
  if (wave == 0) {
    cplx w0 = mk(wr[0], wi[0]), w1 = mk(wr[1], wi[1]);
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      w0 += scr[q * 128 + lk * 16 + li];
      w1 += scr[q * 128 + (lk + 4) * 16 + li];
    }
    // the full W over wave 1's partial (each lane rewrites only the entries it read)
    scr[lk * 16 + li] = w0;
    scr[(lk + 4) * 16 + li] = w1;
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    cplx z0 = mk(0, 0), z1 = mk(0, 0);
#pragma unroll
    for (int l = 0; l < 8; ++l) {
      const cplx wl = scr[l * 16 + li];
      if (ADJ) {  // Z = T^H W
        cfma_cj(z0, Tl[l * 8 + lk], wl);
        cfma_cj(z1, Tl[l * 8 + lk + 4], wl);
      } else {  // Z = T W
        cfma(z0, Tl[lk * 8 + l], wl);
        cfma(z1, Tl[(lk + 4) * 8 + l], wl);
      }
    }
    scr[128 + lk * 16 + li] = z0;
    scr[128 + (lk + 4) * 16 + li] = z1;
  }
  __syncthreads();
  const cplx* Zs = scr + 128;
  const int nrb = (n - k0 + 15) >> 4;
  for (int rb = wave; rb < nrb; rb += 4) {
    const int r0 = k0 + 16 * rb;
    v4d dr, di;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int row = min(r0 + lk + 4 * v, n - 1);
      const cplx x = xc[(size_t)xj * cs + row];
      dr[v] = x.re;
      di[v] = x.im;
    }
#pragma unroll
    for (int ks = 0; ks < 2; ++ks) {
      const int ra = r0 + li;
      const cplx a0 = V[(size_t)(4 * ks + lk) * cs + min(ra, n - 1)];
      const double ar = ra < n ? a0.re : 0.0, am = ra < n ? a0.im : 0.0;
      const cplx z = Zs[(4 * ks + lk) * 16 + li];
      // X -= V Z: re += -Vr Zr + Vi Zi, im += -Vr Zi - Vi Zr
      dr = mfma4(-ar, z.re, dr);
      di = mfma4(-ar, z.im, di);
      dr = mfma4(am, z.im, dr);
      di = mfma4(-am, z.re, di);
    }
    if (li < CB) {
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int row = r0 + lk + 4 * v;
        if (row < n) xc[(size_t)li * cs + row] = mk(dr[v], di[v]);
      }
    }
  }
}

// k_qr_solve on the matrix cores: B <- R^-1 Q^H B with CB columns per workgroup. The panels'
// Q_p^H as in apply_panel_mfma; the back substitution in blocks of 16 rows from the bottom:
// each wave solves the 16 x 16 diagonal block for four columns at once (lane = row + 16 x
// column; the block's inverse diagonal precomputed once), then the rows above take the
// block's contribution as MFMA tiles, R(r, lo:hi) y(lo:hi, :) (K = 16).
template <int CB>
__global__ __launch_bounds__(256) void k_qr_solve_mfma(const ChainDesc* chains, MatRef Ar, MatRef Br, int cs, int nlo, int nhi) {
  int bx, by;
  xcd_block(bx, by);
  const ChainDesc& c = chains[by];
  const int n = c.n;
  const int j0 = bx * CB;
  if (j0 >= n || n <= nlo || n > nhi) return;  // outside this launch's size class
  const int ncol = min(CB, n - j0);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, li = lane & 15, lk = lane >> 4;
  const gcplx* A = gp(matptr(c, Ar));
  gcplx* B = gp(matptr(c, Br));
  extern __shared__ double smem[];
  cplx* xc = reinterpret_cast<cplx*>(smem);  // [CB][cs]
  cplx* V = xc + (size_t)CB * cs;             // [8][cs]; the back substitution's diagonal block
  cplx* scr = V + (size_t)8 * cs;             // panel scratch
  __shared__ cplx Tl[64];
  for (int r = tid; r < n; r += 256) {
    cplx v[CB];
#pragma unroll
    for (int q = 0; q < CB; ++q) v[q] = ld(B + (size_t)(j0 + min(q, ncol - 1)) * n + r);
#pragma unroll
    for (int q = 0; q < CB; ++q) {
      xc[(size_t)q * cs + r].re = v[q].re;  // (a whole-struct copy from a register array is
      xc[(size_t)q * cs + r].im = v[q].im;  // lowered through a stack slot)
    }
  }
  for (int k0 = 0, p = 0; k0 < n; k0 += 8, ++p) {
    const int kb = min(8, n - k0);
    __syncthreads();
    stage_panel<8>(V, A, nullptr, k0, kb, n, tid, 1, cs);  // columns >= kb zero-filled
    if (tid < 64) Tl[tid] = ld(gp(c.qr_Tf) + (size_t)p * 64 + tid);
    __syncthreads();
    apply_panel_mfma<CB, true>(xc, V, Tl, scr, n, cs, k0);
  }
  cplx* Rd = V;        // [16][16]: Rd[kk * 16 + rr] = R(lo + rr, lo + kk)
  cplx* Ri = V + 256;  // [16]: 1 / R(lo + k, lo + k)
  const int xj = min(li, CB - 1);
  for (int hi = n; hi > 0; hi -= 16) {
    const int lo = max(0, hi - 16), bs = hi - lo;
    __syncthreads();
    {
      const int kk = tid >> 4, rr = tid & 15;
      Rd[tid] = (kk < bs && rr < bs) ? ld(A + (size_t)(lo + kk) * n + lo + rr) : mk(0, 0);
      if (tid < 16) Ri[tid] = tid < bs ? cinv(ld(A + (size_t)(lo + tid) * n + lo + tid)) : mk(0, 0);
    }
    __syncthreads();
    {  // wave w: columns w + 4 lk, lane row li
      const int col = wave + 4 * lk;
      cplx* y = xc + (size_t)min(col, CB - 1) * cs;
      cplx zr = li < bs ? y[lo + li] : mk(0, 0);
      for (int k = bs - 1; k >= 0; --k) {
        const cplx yk = shfl_c(zr, (lane & 48) + k) * Ri[k];
        if (li < k) zr = zr - Rd[k * 16 + li] * yk;
        if (li == k) zr = yk;
      }
      if (col < CB && li < bs) y[lo + li] = zr;
    }
    __syncthreads();
    const int nrb = (lo + 15) >> 4;
    for (int rb = wave; rb < nrb; rb += 4) {
      const int r0 = rb * 16;
      v4d dr, di;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const cplx x = xc[(size_t)xj * cs + min(r0 + lk + 4 * v, n - 1)];
        dr[v] = x.re;
        di[v] = x.im;
      }
#pragma unroll
      for (int ks = 0; ks < 4; ++ks) {
        const int kk = 4 * ks + lk, ra = r0 + li;
        const cplx a0 = ld(A + (size_t)(lo + min(kk, bs - 1)) * n + min(ra, n - 1));
        const cplx y0 = xc[(size_t)xj * cs + lo + min(kk, bs - 1)];
        const bool ok = kk < bs && ra < lo;
        const double ar = ok ? a0.re : 0.0, am = ok ? a0.im : 0.0;
        const double yr = kk < bs ? y0.re : 0.0, ym = kk < bs ? y0.im : 0.0;
        // z -= R y
        dr = mfma4(-ar, yr, dr);
        di = mfma4(-ar, ym, di);
        dr = mfma4(am, ym, dr);
        di = mfma4(-am, yr, di);
      }
      if (li < CB) {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int row = r0 + lk + 4 * v;
          if (row < lo) xc[(size_t)li * cs + row] = mk(dr[v], di[v]);
        }
      }
    }
  }
  __syncthreads();
  for (int r = tid; r < n; r += 256) {
    cplx v[CB];
#pragma unroll
    for (int q = 0; q < CB; ++q) v[q] = xc[(size_t)q * cs + r];
#pragma unroll
    for (int q = 0; q < CB; ++q)
      if (q < ncol) st(B + (size_t)(j0 + q) * n + r, v[q]);
  }
}

// k_form_Q_blocked on the matrix cores: CB columns of Q per workgroup from the identity, the
// panels from the last one that reaches them down to the first (Q_p = I - V_p T_p V_p^H).
template <int CB>
__global__ __launch_bounds__(256) void k_form_Q_mfma(const ChainDesc* chains, MatRef Ar, MatRef Qr, int cs, int nlo, int nhi) {
  int bx, by;
  xcd_block(bx, by);
  const ChainDesc& c = chains[by];
  const int n = c.n;
  const int j0 = bx * CB;
  if (j0 >= n || n <= nlo || n > nhi) return;  // outside this launch's size class
  const int ncol = min(CB, n - j0);
  const int tid = threadIdx.x;
  const gcplx* A = gp(matptr(c, Ar));
  gcplx* Q = gp(matptr(c, Qr));
  extern __shared__ double smem[];
  cplx* xc = reinterpret_cast<cplx*>(smem);  // [CB][cs]
  cplx* V = xc + (size_t)CB * cs;             // [8][cs]
  cplx* scr = V + (size_t)8 * cs;
  __shared__ cplx Tl[64];
  for (int e = tid; e < CB * n; e += 256) {
    const int q = e / n, r = e - q * n;
    xc[(size_t)q * cs + r] = mk(r == j0 + q ? 1.0 : 0.0, 0.0);
  }
  const int jmax = j0 + ncol - 1;  // panels starting beyond it leave these columns alone
  for (int p = jmax / 8; p >= 0; --p) {
    const int k0 = p * 8, kb = min(8, n - k0);
    __syncthreads();
    stage_panel<8>(V, A, c.qr_perm, k0, kb, n, tid, 1, cs);  // columns >= kb zero-filled
    if (tid < 64) Tl[tid] = ld(gp(c.qr_Tf) + (size_t)p * 64 + tid);
    __syncthreads();
    apply_panel_mfma<CB, false>(xc, V, Tl, scr, n, cs, k0);
  }
  __syncthreads();
  for (int e = tid; e < ncol * n; e += 256) {
    const int q = e / n, r = e - q * n;
    st(Q + (size_t)j0 * n + e, xc[(size_t)q * cs + r]);
  }
}

// columns per workgroup of the matrix-core solve and Q formation at this size (0: the vector
// kernels): 16 while X, V and the scratch fit the LDS (n <= 400), 8 up to n = 576
inline int mfma_cols(int nmax) { return nmax <= 400 ? 16 : nmax <= 576 ? 8 : 0; }
inline size_t mfma_lds(int nmax, int cb) {
  return ((size_t)(cb + 8) * mfma_ld(nmax) + kMfmaScratch) * sizeof(cplx);
}

// chain counts from which k_form_Q_blocked uses 8 or 16 columns per workgroup (k_qr_solve
// keeps 4: its time per workgroup is the latency of the panel sequence, and wider blocks
// measured slower)
constexpr int kWideColChains = 8;

// default panel width: 8 (the register-resident panel keeps nb x rows-per-thread values per
// thread; at nb = 16 the per-column chain is longer and the kernel spills)
int qr_panel_width(int nmax) {
  (void)nmax;
  return 8;
}

template <int NB, int RQ>
hipError_t launch_qr_blocked_nb(hipStream_t s, const ChainDesc* chains, int nchains, int nmax, MatRef A, VecRef D,
                                int prepivot, int wg, const int* bmap, int nblocks) {
  int ps = mfma_ld(nmax);
  const size_t lds = ((size_t)NB * ps + 4 * 128) * sizeof(cplx);
  {
    hipError_t e = ensure_dynamic_lds((const void*)k_qr_blocked<NB, RQ>, lds);
    if (e != hipSuccess) return e;
  }
  void* args[] = {(void*)&chains, (void*)&A, (void*)&D, (void*)&prepivot, (void*)&ps, (void*)&bmap};
  const dim3 grid = bmap ? dim3(nblocks, 1) : dim3(wg, nchains);
  hipError_t e = hipLaunchCooperativeKernel((const void*)k_qr_blocked<NB, RQ>, grid, dim3(256), args, lds, s);
  if (e != hipSuccess) (void)hipGetLastError();
  return e;
}

hipError_t launch_qr_blocked(hipStream_t s, const ChainDesc* chains, int nchains, int nmax, MatRef A, VecRef D,
                             int prepivot, int wg, int nb, const int* bmap, int nblocks) {
  if (nmax > kQrMaxN) return hipErrorInvalidValue;
  if (nb == 0) nb = qr_panel_width(nmax);
  if (nb == 16 && nmax <= 512)
    return launch_qr_blocked_nb<16, 2>(s, chains, nchains, nmax, A, D, prepivot, wg, bmap, nblocks);
  if (nb == 8 && nmax <= 512)
    return launch_qr_blocked_nb<8, 2>(s, chains, nchains, nmax, A, D, prepivot, wg, bmap, nblocks);
  // three rows per thread up to n = 768 (four spill the register-resident panel)
  if (nb == 8 && nmax <= 768)
    return launch_qr_blocked_nb<8, 3>(s, chains, nchains, nmax, A, D, prepivot, wg, bmap, nblocks);
  if (nb == 8) return launch_qr_blocked_nb<8, 4>(s, chains, nchains, nmax, A, D, prepivot, wg, bmap, nblocks);
  return hipErrorInvalidValue;
}

template <int NB, int CB>
hipError_t launch_qr_solve_nb(hipStream_t s, const ChainDesc* chains, int nchains, int nmax, MatRef A, MatRef B) {
  const size_t lds = (size_t)(CB + NB) * nmax * sizeof(cplx);
  {
    hipError_t e = ensure_dynamic_lds((const void*)k_qr_solve<NB, CB>, lds);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL((k_qr_solve<NB, CB>), dim3((nmax + CB - 1) / CB, nchains), dim3(256), lds, s, chains, A, B);
  return hipGetLastError();
}

template <int NB, int CB>
hipError_t launch_form_Q_blocked_nb(hipStream_t s, const ChainDesc* chains, int nchains, int nmax, MatRef A,
                                    MatRef Q) {
  const size_t lds = (size_t)(CB + NB) * nmax * sizeof(cplx);
  {
    hipError_t e = ensure_dynamic_lds((const void*)k_form_Q_blocked<NB, CB>, lds);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL((k_form_Q_blocked<NB, CB>), dim3((nmax + CB - 1) / CB, nchains), dim3(256), lds, s, chains, A,
                     Q);
  return hipGetLastError();
}

template <int CB>
hipError_t launch_mfma_kernel(const void* fn, hipStream_t s, int nchains, int nmax, const ChainDesc* chains, MatRef A,
                              MatRef X, int nlo, int nhi) {
  const size_t lds = mfma_lds(nmax, CB);
  {
    hipError_t e = ensure_dynamic_lds(fn, lds);
    if (e != hipSuccess) return e;
  }
  int cs = mfma_ld(nmax);
  void* args[] = {(void*)&chains, (void*)&A, (void*)&X, (void*)&cs, (void*)&nlo, (void*)&nhi};
  hipError_t e = hipLaunchKernel(fn, dim3((nmax + CB - 1) / CB, nchains), dim3(256), args, lds, s);
  if (e != hipSuccess) (void)hipGetLastError();
  return e;
}

// the matrix-core solve / Q formation over chains up to nmax: chains with n <= 400 take 16
// columns per workgroup, larger ones 8 (two launches when the sizes are mixed; the chains of
// the other class return at once)
template <class F16, class F8>
hipError_t launch_mfma_classes(F16 k16, F8 k8, hipStream_t s, int nchains, int nmax, const ChainDesc* chains, MatRef A,
                               MatRef X) {
  if (nmax <= 400) return launch_mfma_kernel<16>((const void*)k16, s, nchains, nmax, chains, A, X, 0, nmax);
  hipError_t e = launch_mfma_kernel<16>((const void*)k16, s, nchains, 400, chains, A, X, 0, 400);
  if (e != hipSuccess) return e;
  return launch_mfma_kernel<8>((const void*)k8, s, nchains, nmax, chains, A, X, 400, nmax);
}

hipError_t launch_form_Q_blocked(hipStream_t s, const ChainDesc* chains, int nchains, int nmax, MatRef A, MatRef Q,
                                 int nb) {
  if (nmax > kQrMaxN || nmax < 16) return hipErrorInvalidValue;
  if (nb == 0) nb = qr_panel_width(nmax);
  if (nb == 16 && nmax <= 512) return launch_form_Q_blocked_nb<16, 4>(s, chains, nchains, nmax, A, Q);
  if (nb != 8) return hipErrorInvalidValue;
  if (nchains >= kWideColChains) {
    const int cb = mfma_cols(nmax);
    if (cb) return launch_mfma_classes(k_form_Q_mfma<16>, k_form_Q_mfma<8>, s, nchains, nmax, chains, A, Q);
  }
  // many chains fill the chip with fewer, wider column blocks: each reads the reflectors once
  if (nchains >= kWideColChains && nmax <= 400) return launch_form_Q_blocked_nb<8, 16>(s, chains, nchains, nmax, A, Q);
  if (nchains >= kWideColChains && nmax <= 576) return launch_form_Q_blocked_nb<8, 8>(s, chains, nchains, nmax, A, Q);
  if (nmax <= 640) return launch_form_Q_blocked_nb<8, 4>(s, chains, nchains, nmax, A, Q);
  return launch_form_Q_blocked_nb<8, 1>(s, chains, nchains, nmax, A, Q);
}

hipError_t launch_qr_solve(hipStream_t s, const ChainDesc* chains, int nchains, int nmax, MatRef A, MatRef B, int nb) {
  if (nmax > kQrMaxN || nmax < 16) return hipErrorInvalidValue;
  if (nb == 0) nb = qr_panel_width(nmax);
  if (nb == 16 && nmax <= 512) return launch_qr_solve_nb<16, 4>(s, chains, nchains, nmax, A, B);
  if (nb != 8) return hipErrorInvalidValue;
  if (nchains >= kWideColChains) {
    const int cb = mfma_cols(nmax);
    if (cb) return launch_mfma_classes(k_qr_solve_mfma<16>, k_qr_solve_mfma<8>, s, nchains, nmax, chains, A, B);
  }
  if (nmax <= 640) return launch_qr_solve_nb<8, 4>(s, chains, nchains, nmax, A, B);
  return launch_qr_solve_nb<8, 1>(s, chains, nchains, nmax, A, B);
}

}  // namespace dqmc
