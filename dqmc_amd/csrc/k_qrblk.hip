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
// and factors panel p+1 in LDS (zlarfg per column, applied to the rest of the panel, then
// zlarft) while the other workgroups apply Q_p^H to the columns beyond (one wave per
// column); one group barrier per panel.
#include "chain.hpp"

namespace dqmc {

constexpr int kQrMaxN = 1024;
constexpr int kQrMaxNb = 16;

__device__ inline cplx wsum(cplx v) { return wave_sum(v); }
__device__ inline cplx shfl_c(cplx v, int src) { return mk(__shfl(v.re, src, 64), __shfl(v.im, src, 64)); }
__device__ inline cplx read_lane_c(cplx v, int lane) { return mk(read_lane(v.re, lane), read_lane(v.im, lane)); }

// Panel staging between A (global) and P (LDS, [NB][n], rows k0..n-1 of each column);
// logical column k is physical column perm[k] (perm == nullptr: the identity).
// Each thread moves one row of all NB columns per step: the NB loads are issued before
// any store, so memory latency is paid once per row rather than once per element.
// mode 0: P <- A as is; mode 1: P <- unit lower trapezoid (V with its implicit ones and
// zeros above the diagonal), columns t >= kb zero-filled.
template <int NB>
__device__ inline void stage_panel(cplx* P, const cplx* A, const int* perm, int k0, int kb, int n, int tid,
                                   int mode) {
  const cplx* src[NB];
#pragma unroll
  for (int t = 0; t < NB; ++t) {
    const int k = k0 + min(t, kb - 1);
    src[t] = A + (size_t)(perm ? perm[k] : k) * n;
  }
  for (int r = k0 + tid; r < n; r += 256) {
    cplx v[NB];
#pragma unroll
    for (int t = 0; t < NB; ++t) v[t] = src[t][r];
#pragma unroll
    for (int t = 0; t < NB; ++t) {
      cplx val = v[t];
      if (mode == 1) val = (t >= kb || r < k0 + t) ? mk(0, 0) : (r == k0 + t) ? mk(1, 0) : val;
      P[(size_t)t * n + r] = val;
    }
  }
}

template <int NB>
__device__ inline void unstage_panel(cplx* A, const cplx* P, const int* perm, int k0, int kb, int n, int tid) {
  cplx* dst[NB];
#pragma unroll
  for (int t = 0; t < NB; ++t) dst[t] = A + (size_t)perm[k0 + min(t, kb - 1)] * n;
  for (int r = k0 + tid; r < n; r += 256) {
    cplx v[NB];
#pragma unroll
    for (int t = 0; t < NB; ++t) v[t] = P[(size_t)min(t, kb - 1) * n + r];
#pragma unroll
    for (int t = 0; t < NB; ++t) {  // t >= kb rewrites column kb-1 with its own values
      dst[t][r].re = v[t].re;  // (a whole-struct copy here is lowered through a stack slot)
      dst[t][r].im = v[t].im;
    }
  }
}

// Inner loops below are written without data-dependent branches around loads: a guard
// around each LDS read makes the compiler wait for it before the next one, which
// serialises the loop on LDS latency. Out-of-range rows and columns are read from
// clamped in-bounds addresses and their contributions zeroed; masked-off stores go to a
// per-thread sink slot.
template <int NB>
__global__ __launch_bounds__(256) void k_qr_blocked(const ChainDesc* chains, MatRef Ar, VecRef Dr, int prepivot) {
  constexpr int KC = (NB + 3) / 4;  // panel columns per wave in the per-column update
  constexpr int QN = (NB == 16 ? 512 : kQrMaxN) / 64;  // row blocks of a column (NB = 16 only for n <= 512)
  const ChainDesc& c = chains[blockIdx.y];
  const int n = c.n;
  const int W = gridDim.x, w = blockIdx.x;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, nwaves = blockDim.x >> 6;
  cplx* A = matptr(c, Ar);
  double* D = vecptr(c, Dr);
  GroupSync g = chain_group(c);

  extern __shared__ double smem[];
  cplx* P = reinterpret_cast<cplx*>(smem);  // [NB][n] panel (rows k0..n-1 used) / staged V
  cplx* sink = P + (size_t)NB * n;           // [256] one slot per thread
  __shared__ int perm[kQrMaxN], lpos[kQrMaxN];
  __shared__ cplx Tl[NB * NB];  // Tl[j * NB + t] = T(j, t)
  __shared__ cplx taul[NB];
  __shared__ double betal[NB];

  // ---- column order
  if (prepivot) {
    for (int col = w + W * wave; col < n; col += W * nwaves) {
      const cplx* a = A + (size_t)col * n;
      double acc = 0.0;
      for (int k = lane; k < n; k += 64) acc += abs2(a[k]);
      acc = wave_sum(acc);
      if (lane == 0) c.qr_vn[col] = sqrt(acc);
    }
    group_sync(g);
    if (group_failed(g)) return;
    double* vn = reinterpret_cast<double*>(P);  // n doubles, before the panel buffer is used
    for (int k = tid; k < n; k += blockDim.x) vn[k] = c.qr_vn[k];
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

  // ---- the panel factorisation of panel p (workgroup 0 alone, in LDS)
  auto factor_panel = [&](int k0, int p) {
    const int kb = min(NB, n - k0);
    cplx* Tg = c.qr_Tf + (size_t)p * NB * NB;
    stage_panel<NB>(P, A, perm, k0, kb, n, tid, 0);
    __syncthreads();
    // per column: wave 0 forms the reflector (zlarfg) alone, one barrier publishes it, every
    // wave applies it to its share of the remaining panel columns, one barrier
    for (int t = 0; t < kb; ++t) {
      const int i = k0 + t;
      cplx* x = P + (size_t)t * n;
      if (wave == 0) {
        double part = 0.0;
        for (int kq = i + 1 + lane; kq < n; kq += 256) {
          cplx v[4];
#pragma unroll
          for (int q = 0; q < 4; ++q) v[q] = x[min(kq + 64 * q, n - 1)];
#pragma unroll
          for (int q = 0; q < 4; ++q) part += (kq + 64 * q < n) ? abs2(v[q]) : 0.0;
        }
        const double xnorm = sqrt(wave_sum(part));
        const cplx alpha = x[i];
        cplx tau;
        double beta;
        // zlarfg: H^H (alpha, x) = (beta, 0), v = (1, x / (alpha - beta))
        if (xnorm == 0.0 && alpha.im == 0.0) {
          tau = mk(0, 0);
          beta = alpha.re;
        } else {
          beta = -copysign(sqrt(alpha.re * alpha.re + alpha.im * alpha.im + xnorm * xnorm), alpha.re);
          tau = mk((beta - alpha.re) / beta, -alpha.im / beta);
          const cplx sc = cinv(alpha - mk(beta, 0));
          for (int kq = i + 1 + lane; kq < n; kq += 256) {
            cplx v[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) v[q] = x[min(kq + 64 * q, n - 1)];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
              const int k = kq + 64 * q;
              *(k < n ? x + k : sink + tid) = v[q] * sc;
            }
          }
        }
        if (lane == 0) {  // global copies are written after the panel: a store here
          taul[t] = tau;   // would make every later __syncthreads wait for it
          betal[t] = beta;
        }
      }
      __syncthreads();
      // H^H = I - conj(tau) v v^H on the remaining panel columns: wave w takes columns
      // c0 = t+1+w, c0+4, ... (nu of them, wave-uniform); their loads and sums overlap
      const int c0 = t + 1 + wave;
      const int nu = c0 < kb ? (kb - c0 + 3) >> 2 : 0;
      if (nu > 0) {
        const cplx ctau = conj(taul[t]);
        cplx* col[KC];
#pragma unroll
        for (int u = 0; u < KC; ++u) col[u] = (u < nu) ? P + (size_t)(c0 + 4 * u) * n : x;
        cplx acc[KC];
#pragma unroll
        for (int u = 0; u < KC; ++u) acc[u] = mk(0, 0);
        for (int kq = i + lane; kq < n; kq += 128) {
          cplx vk[2], pv[2][KC];
#pragma unroll
          for (int q = 0; q < 2; ++q) {
            const int kk = min(kq + 64 * q, n - 1);
            vk[q] = x[kk];
#pragma unroll
            for (int u = 0; u < KC; ++u) pv[q][u] = col[u][kk];
          }
#pragma unroll
          for (int q = 0; q < 2; ++q) {
            const int k = kq + 64 * q;
            const cplx v = (k == i) ? mk(1, 0) : (k < n) ? vk[q] : mk(0, 0);
#pragma unroll
            for (int u = 0; u < KC; ++u) cfma_cj(acc[u], v, pv[q][u]);
          }
        }
        cplx f[KC];
#pragma unroll
        for (int u = 0; u < KC; ++u) f[u] = (u < nu) ? ctau * wsum(acc[u]) : mk(0, 0);
        for (int kq = i + lane; kq < n; kq += 128) {
          cplx vk[2], pv[2][KC];
#pragma unroll
          for (int q = 0; q < 2; ++q) {
            const int kk = min(kq + 64 * q, n - 1);
            vk[q] = x[kk];
#pragma unroll
            for (int u = 0; u < KC; ++u) pv[q][u] = col[u][kk];
          }
#pragma unroll
          for (int q = 0; q < 2; ++q) {
            const int k = kq + 64 * q;
            const cplx v = (k == i) ? mk(1, 0) : vk[q];
#pragma unroll
            for (int u = 0; u < KC; ++u) *((u < nu && k < n) ? col[u] + k : sink + tid) = pv[q][u] - f[u] * v;
          }
        }
      }
      __syncthreads();
    }
    // R's diagonal goes where the reflectors' implicit ones were (never read again)
    if (tid < kb) {
      P[(size_t)tid * n + k0 + tid] = mk(betal[tid], 0);
      D[k0 + tid] = fabs(betal[tid]);
      c.qr_tau[k0 + tid] = taul[tid];
    }
    // ---- zlarft (forward, columnwise): T(t,t) = tau_t, T(0:t, t) = -tau_t T(0:t, 0:t) V(:, 0:t)^H v_t
    // first the inner products S(j, t) = V(:, j)^H v_t, j < t, into Tl(j, t); wave w takes
    // t = 1+w, 5+w, ...; the products of one t share v_t's loads. Columns j >= t are read
    // too (in bounds) and their sums dropped.
    for (int t = 1 + wave; t < kb; t += nwaves) {
      const cplx* vt = P + (size_t)t * n;
      const int it = k0 + t;
      cplx acc[NB];
#pragma unroll
      for (int j = 0; j < NB; ++j) acc[j] = mk(0, 0);
      for (int k = it + lane; k < n; k += 64) {
        const cplx b = (k == it) ? mk(1, 0) : vt[k];
        cplx pj[NB];
#pragma unroll
        for (int j = 0; j < NB; ++j) pj[j] = P[(size_t)j * n + k];  // rows >= it: below v_j's diagonal
#pragma unroll
        for (int j = 0; j < NB; ++j) cfma_cj(acc[j], pj[j], b);
      }
#pragma unroll
      for (int j = 0; j < NB; ++j)
        if (j < t) {
          const cplx sj = wsum(acc[j]);
          if (lane == 0) Tl[j * NB + t] = sj;
        }
    }
    __syncthreads();
    // the triangle, by wave 0: lane j keeps row j of T in registers; column t needs
    // sum_{l<t} T(j, l) S(l, t), where T(j, l) = 0 for j > l
    if (wave == 0) {
      cplx trow[NB];
#pragma unroll
      for (int l = 0; l < NB; ++l) trow[l] = mk(0, 0);
#pragma unroll
      for (int t = 0; t < NB; ++t) {
        if (t < kb) {
          cplx acc = mk(0, 0);
#pragma unroll
          for (int l = 0; l < t; ++l) cfma(acc, trow[l], Tl[l * NB + t]);
          const cplx tt = taul[t];
          trow[t] = (lane < t) ? -(tt * acc) : (lane == t) ? tt : mk(0, 0);
        }
      }
      if (lane < NB) {
#pragma unroll
        for (int t = 0; t < NB; ++t) {
          Tl[lane * NB + t] = trow[t];
          Tg[lane * NB + t] = trow[t];
        }
      }
    }
    __syncthreads();
    unstage_panel<NB>(A, P, perm, k0, kb, n, tid);
  };
  // ---- a <- (I - V T^H V^H) a for one column, with panel V (from row k0) staged in P and T in Tl
  auto apply_col = [&](int col, int k0) {
    cplx* a = A + (size_t)col * n;
    // the column's rows k0.. in registers (all loads in flight together)
    cplx av[QN];
#pragma unroll
    for (int q = 0; q < QN; ++q) {
      const int k = k0 + lane + 64 * q;
      av[q] = k < n ? a[k] : mk(0, 0);
    }
    cplx wv[NB];
#pragma unroll
    for (int j = 0; j < NB; ++j) wv[j] = mk(0, 0);
#pragma unroll
    for (int q = 0; q < QN; ++q) {
      const int k = k0 + lane + 64 * q;
      if (k0 + 64 * q < n) {  // uniform
        const int kk = min(k, n - 1);
        cplx pj[NB];
#pragma unroll
        for (int j = 0; j < NB; ++j) pj[j] = P[(size_t)j * n + kk];
#pragma unroll
        for (int j = 0; j < NB; ++j) cfma_cj(wv[j], pj[j], av[q]);  // av = 0 past row n-1
      }
    }
#pragma unroll
    for (int j = 0; j < NB; ++j) wv[j] = wsum(wv[j]);
    // z = T^H w: lane j < NB forms z_j, then every lane reads all NB of them
    const int jl = lane & (NB - 1);
    cplx zl = mk(0, 0);
#pragma unroll
    for (int l = 0; l < NB; ++l) cfma_cj(zl, Tl[l * NB + jl], wv[l]);
    cplx z[NB];
#pragma unroll
    for (int j = 0; j < NB; ++j) z[j] = read_lane_c(zl, j);
#pragma unroll
    for (int q = 0; q < QN; ++q) {
      const int k = k0 + lane + 64 * q;
      if (k0 + 64 * q < n) {
        const int kk = min(k, n - 1);
        cplx pj[NB];
#pragma unroll
        for (int j = 0; j < NB; ++j) pj[j] = P[(size_t)j * n + kk];
        cplx ak = av[q];
#pragma unroll
        for (int j = 0; j < NB; ++j) ak = ak - pj[j] * z[j];
        if (k < n) a[k] = ak;
      }
    }
  };

  // One-panel look-ahead: while every other workgroup applies panel p to the columns beyond
  // panel p+1, workgroup 0 applies it to panel p+1's columns and factors panel p+1, so each
  // panel costs one group barrier and its factorisation overlaps the trailing update.
  if (w == 0) factor_panel(0, 0);
  group_sync(g);
  if (group_failed(g)) return;
  for (int k0 = 0, p = 0; k0 + NB < n; k0 += NB, ++p) {  // panel p (full width) is factored
    const int k1 = k0 + NB, kb1 = min(NB, n - k1);
    stage_panel<NB>(P, A, perm, k0, NB, n, tid, 1);
    for (int e = tid; e < NB * NB; e += blockDim.x) Tl[e] = c.qr_Tf[(size_t)p * NB * NB + e];
    __syncthreads();
    if (w == 0) {
      for (int u = wave; u < kb1; u += nwaves) apply_col(perm[k1 + u], k0);
      if (W == 1)  // nobody else to take the trailing columns
        for (int col = wave; col < n; col += nwaves)
          if (lpos[col] >= k1 + NB) apply_col(col, k0);
      __syncthreads();
      factor_panel(k1, p + 1);
    } else {
      for (int col = (w - 1) + (W - 1) * wave; col < n; col += (W - 1) * nwaves)
        if (lpos[col] >= k1 + NB) apply_col(col, k0);
    }
    group_sync(g);
    if (group_failed(g)) return;
  }
  if (w == 0) {
    for (int k = tid; k < n; k += blockDim.x) {
      c.qr_perm[k] = perm[k];
      c.qr_lpos[k] = lpos[k];
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
  constexpr int RS = 4 / CB;
  int bx, by;
  xcd_block(bx, by);
  const ChainDesc& c = chains[by];
  const int n = c.n;
  const int j0 = bx * CB;
  if (j0 >= n) return;
  const int ncol = min(CB, n - j0);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int jl = wave % CB, rs = wave / CB;
  const cplx* A = matptr(c, Ar);
  cplx* B = matptr(c, Br);
  extern __shared__ double smem[];
  cplx* bc = reinterpret_cast<cplx*>(smem);  // [CB][n]
  cplx* V = bc + (size_t)CB * n;              // [NB][n]
  __shared__ cplx Tl[NB * NB];
  __shared__ cplx Rd[16 * 16];
  __shared__ cplx part[RS > 1 ? 4 * NB : 1];
  for (int r = tid; r < n; r += 256) {
    cplx v[CB];
#pragma unroll
    for (int q = 0; q < CB; ++q) {
      const cplx* src = B + (size_t)(j0 + min(q, ncol - 1)) * n + r;
      v[q] = mk(src->re, src->im);  // (a whole-struct copy here is lowered through a stack slot)
    }
#pragma unroll
    for (int q = 0; q < CB; ++q) bc[(size_t)q * n + r] = v[q];
  }
  const bool active = jl < ncol;  // wave-uniform
  cplx* a = bc + (size_t)jl * n;
  for (int k0 = 0, p = 0; k0 < n; k0 += NB, ++p) {
    const int kb = min(NB, n - k0);
    __syncthreads();
    stage_panel<NB>(V, A, nullptr, k0, kb, n, tid, 1);  // columns >= kb zero-filled
    for (int e = tid; e < NB * NB; e += blockDim.x) Tl[e] = c.qr_Tf[(size_t)p * NB * NB + e];
    __syncthreads();
    // w = V^H a over this wave's row slice (rows k0 + lane + 64 (rs + RS m))
    cplx wv[NB];
#pragma unroll
    for (int j = 0; j < NB; ++j) wv[j] = mk(0, 0);
    for (int k = k0 + lane + 64 * rs; k < n; k += 64 * RS) {
      const cplx ak = a[k];
      cplx pj[NB];
#pragma unroll
      for (int j = 0; j < NB; ++j) pj[j] = V[(size_t)j * n + k];
#pragma unroll
      for (int j = 0; j < NB; ++j) cfma_cj(wv[j], pj[j], ak);
    }
#pragma unroll
    for (int j = 0; j < NB; ++j) wv[j] = wsum(wv[j]);
    if (RS > 1) {
      if (lane == 0) {
#pragma unroll
        for (int j = 0; j < NB; ++j) part[wave * NB + j] = wv[j];
      }
      __syncthreads();
#pragma unroll
      for (int j = 0; j < NB; ++j) {
        cplx sj = mk(0, 0);
#pragma unroll
        for (int q = 0; q < RS; ++q) sj += part[(jl + CB * q) * NB + j];
        wv[j] = sj;
      }
    }
    // z = T^H w: lane j < NB forms z_j, then every lane reads all of them
    const int jz = lane & (NB - 1);
    cplx zl = mk(0, 0);
#pragma unroll
    for (int l = 0; l < NB; ++l) cfma_cj(zl, Tl[l * NB + jz], wv[l]);
    cplx z[NB];
#pragma unroll
    for (int j = 0; j < NB; ++j) z[j] = read_lane_c(zl, j);
    if (active) {
      for (int k = k0 + lane + 64 * rs; k < n; k += 64 * RS) {
        cplx ak = a[k];
        cplx pj[NB];
#pragma unroll
        for (int j = 0; j < NB; ++j) pj[j] = V[(size_t)j * n + k];
#pragma unroll
        for (int j = 0; j < NB; ++j) ak = ak - pj[j] * z[j];
        a[k] = ak;
      }
    }
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
      Rd[tid] = (kk < bs && rr < bs) ? A[(size_t)(lo + kk) * n + lo + rr] : mk(0, 0);  // Rd[kk*16+rr] = R(lo+rr, lo+kk)
    }
    __syncthreads();
    if (wave < ncol) {
      cplx* y = bc + (size_t)wave * n;
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
      for (int k = 0; k < 16; ++k) rv[k] = A[(size_t)(lo + min(k, bs - 1)) * n + r];
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
        cplx* d = B + (size_t)(j0 + q) * n + r;
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
  constexpr int RS = 4 / CB;
  int bx, by;
  xcd_block(bx, by);
  const ChainDesc& c = chains[by];
  const int n = c.n;
  const int j0 = bx * CB;
  if (j0 >= n) return;
  const int ncol = min(CB, n - j0);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int jl = wave % CB, rs = wave / CB;
  const cplx* A = matptr(c, Ar);
  cplx* Q = matptr(c, Qr);
  extern __shared__ double smem[];
  cplx* xc = reinterpret_cast<cplx*>(smem);  // [CB][n]
  cplx* V = xc + (size_t)CB * n;              // [NB][n]
  __shared__ cplx Tl[NB * NB];
  __shared__ cplx part[RS > 1 ? 4 * NB : 1];
  for (int r = tid; r < n; r += 256)
#pragma unroll
    for (int q = 0; q < CB; ++q) {
      xc[(size_t)q * n + r].re = (r == j0 + q) ? 1.0 : 0.0;
      xc[(size_t)q * n + r].im = 0.0;
    }
  const bool active = jl < ncol;  // wave-uniform
  cplx* a = xc + (size_t)jl * n;
  const int jmax = j0 + ncol - 1;  // panels starting beyond it leave these columns alone
  for (int p = jmax / NB; p >= 0; --p) {
    const int k0 = p * NB, kb = min(NB, n - k0);
    __syncthreads();
    stage_panel<NB>(V, A, c.qr_perm, k0, kb, n, tid, 1);  // columns >= kb zero-filled
    for (int e = tid; e < NB * NB; e += blockDim.x) Tl[e] = c.qr_Tf[(size_t)p * NB * NB + e];
    __syncthreads();
    cplx wv[NB];
#pragma unroll
    for (int j = 0; j < NB; ++j) wv[j] = mk(0, 0);
    for (int k = k0 + lane + 64 * rs; k < n; k += 64 * RS) {
      const cplx ak = a[k];
      cplx pj[NB];
#pragma unroll
      for (int j = 0; j < NB; ++j) pj[j] = V[(size_t)j * n + k];
#pragma unroll
      for (int j = 0; j < NB; ++j) cfma_cj(wv[j], pj[j], ak);
    }
#pragma unroll
    for (int j = 0; j < NB; ++j) wv[j] = wsum(wv[j]);
    if (RS > 1) {
      if (lane == 0) {
#pragma unroll
        for (int j = 0; j < NB; ++j) part[wave * NB + j] = wv[j];
      }
      __syncthreads();
#pragma unroll
      for (int j = 0; j < NB; ++j) {
        cplx sj = mk(0, 0);
#pragma unroll
        for (int q = 0; q < RS; ++q) sj += part[(jl + CB * q) * NB + j];
        wv[j] = sj;
      }
    }
    // z = T w: lane j < NB forms z_j, then every lane reads all of them
    const int jz = lane & (NB - 1);
    cplx zl = mk(0, 0);
#pragma unroll
    for (int l = 0; l < NB; ++l) cfma(zl, Tl[jz * NB + l], wv[l]);
    cplx z[NB];
#pragma unroll
    for (int j = 0; j < NB; ++j) z[j] = read_lane_c(zl, j);
    if (active) {
      for (int k = k0 + lane + 64 * rs; k < n; k += 64 * RS) {
        cplx ak = a[k];
        cplx pj[NB];
#pragma unroll
        for (int j = 0; j < NB; ++j) pj[j] = V[(size_t)j * n + k];
#pragma unroll
        for (int j = 0; j < NB; ++j) ak = ak - pj[j] * z[j];
        a[k] = ak;
      }
    }
  }
  __syncthreads();
  for (int r = tid; r < n; r += 256) {
    cplx v[CB];
#pragma unroll
    for (int q = 0; q < CB; ++q) v[q] = xc[(size_t)q * n + r];
#pragma unroll
    for (int q = 0; q < CB; ++q)
      if (q < ncol) {
        cplx* d = Q + (size_t)(j0 + q) * n + r;
        d->re = v[q].re;
        d->im = v[q].im;
      }
  }
}

int qr_panel_width(int nmax) { return nmax <= 512 ? 16 : 8; }

template <int NB>
hipError_t launch_qr_blocked_nb(hipStream_t s, const ChainDesc* chains, int nchains, int nmax, MatRef A, VecRef D,
                                int prepivot, int wg) {
  const size_t lds = ((size_t)NB * nmax + 256) * sizeof(cplx);
  static size_t attr = 0;
  if (lds > 65536 && lds > attr) {
    hipError_t e =
        hipFuncSetAttribute((const void*)k_qr_blocked<NB>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    attr = lds;
  }
  void* args[] = {(void*)&chains, (void*)&A, (void*)&D, (void*)&prepivot};
  hipError_t e = hipLaunchCooperativeKernel((const void*)k_qr_blocked<NB>, dim3(wg, nchains), dim3(256), args, lds, s);
  if (e != hipSuccess) (void)hipGetLastError();
  return e;
}

hipError_t launch_qr_blocked(hipStream_t s, const ChainDesc* chains, int nchains, int nmax, MatRef A, VecRef D,
                             int prepivot, int wg, int nb) {
  if (nmax > kQrMaxN) return hipErrorInvalidValue;
  if (nb == 0) nb = qr_panel_width(nmax);
  if (nb == 16 && nmax <= 512) return launch_qr_blocked_nb<16>(s, chains, nchains, nmax, A, D, prepivot, wg);
  if (nb == 8) return launch_qr_blocked_nb<8>(s, chains, nchains, nmax, A, D, prepivot, wg);
  return hipErrorInvalidValue;
}

template <int NB, int CB>
hipError_t launch_qr_solve_nb(hipStream_t s, const ChainDesc* chains, int nchains, int nmax, MatRef A, MatRef B) {
  const size_t lds = (size_t)(CB + NB) * nmax * sizeof(cplx);
  static size_t attr = 0;
  if (lds > 65536 && lds > attr) {
    hipError_t e =
        hipFuncSetAttribute((const void*)k_qr_solve<NB, CB>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    attr = lds;
  }
  hipLaunchKernelGGL((k_qr_solve<NB, CB>), dim3((nmax + CB - 1) / CB, nchains), dim3(256), lds, s, chains, A, B);
  return hipGetLastError();
}

template <int NB, int CB>
hipError_t launch_form_Q_blocked_nb(hipStream_t s, const ChainDesc* chains, int nchains, int nmax, MatRef A,
                                    MatRef Q) {
  const size_t lds = (size_t)(CB + NB) * nmax * sizeof(cplx);
  static size_t attr = 0;
  if (lds > 65536 && lds > attr) {
    hipError_t e = hipFuncSetAttribute((const void*)k_form_Q_blocked<NB, CB>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    attr = lds;
  }
  hipLaunchKernelGGL((k_form_Q_blocked<NB, CB>), dim3((nmax + CB - 1) / CB, nchains), dim3(256), lds, s, chains, A,
                     Q);
  return hipGetLastError();
}

hipError_t launch_form_Q_blocked(hipStream_t s, const ChainDesc* chains, int nchains, int nmax, MatRef A, MatRef Q,
                                 int nb) {
  if (nmax > kQrMaxN || nmax < 16) return hipErrorInvalidValue;
  if (nb == 0) nb = qr_panel_width(nmax);
  if (nb == 16 && nmax <= 512) return launch_form_Q_blocked_nb<16, 4>(s, chains, nchains, nmax, A, Q);
  if (nb != 8) return hipErrorInvalidValue;
  if (nmax <= 640) return launch_form_Q_blocked_nb<8, 4>(s, chains, nchains, nmax, A, Q);
  return launch_form_Q_blocked_nb<8, 1>(s, chains, nchains, nmax, A, Q);
}

hipError_t launch_qr_solve(hipStream_t s, const ChainDesc* chains, int nchains, int nmax, MatRef A, MatRef B, int nb) {
  if (nmax > kQrMaxN || nmax < 16) return hipErrorInvalidValue;
  if (nb == 0) nb = qr_panel_width(nmax);
  if (nb == 16 && nmax <= 512) return launch_qr_solve_nb<16, 4>(s, chains, nchains, nmax, A, B);
  if (nb != 8) return hipErrorInvalidValue;
  if (nmax <= 640) return launch_qr_solve_nb<8, 4>(s, chains, nchains, nmax, A, B);
  return launch_qr_solve_nb<8, 1>(s, chains, nchains, nmax, A, B);
}

}  // namespace dqmc
