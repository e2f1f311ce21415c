// Blocked Householder QR on a chain group, and the solve that uses it.
//
// k_qr_blocked factors A P0 = Q R panel by panel (compact WY, Q_p = I - V_p T_p V_p^H,
// LAPACK zgeqrf/zlarft conventions). P0 is either the identity or the pre-pivoting
// order: columns sorted once by decreasing initial norm (ties by index). For the
// column-graded products DQMC stabilises, (B...B U) diag(D), QR with pre-pivoting is
// as stable as column-pivoted QR (Bai, Lee, Li, Xu, "Stable solutions of linear systems
// involving long chain of matrix multiplications", Linear Algebra Appl. 435 (2011)),
// while needing only two group barriers per panel instead of one per column.
//
// The output uses the layout of k_qrp (columns never move: logical column k lives in
// physical column qr_perm[k]; reflector k below the diagonal of that column, R on and
// above it; qr_tau[k]; D[k] = |Re R_kk|), so k_form_T / k_form_Q form U and T of
// decompose_udt (linalg.jl:20-39) unchanged. The T factors of all panels are kept in
// qr_Tf for k_qr_solve.
//
// Per panel: workgroup 0 factors the panel in LDS (zlarfg per column, applied to the
// rest of the panel, then zlarft), writes it back, group barrier; every workgroup
// applies Q_p^H to its own trailing columns (one wave per column), group barrier.
#include "chain.hpp"

namespace dqmc {

constexpr int kQrMaxN = 1024;
constexpr int kQrMaxNb = 16;

__device__ inline cplx wsum(cplx v) { return wave_sum(v); }
__device__ inline cplx shfl_c(cplx v, int src) { return mk(__shfl(v.re, src, 64), __shfl(v.im, src, 64)); }

__global__ __launch_bounds__(256) void k_qr_blocked(const ChainDesc* chains, MatRef Ar, VecRef Dr, int prepivot,
                                                    int nb) {
  const ChainDesc& c = chains[blockIdx.y];
  const int n = c.n;
  const int W = gridDim.x, w = blockIdx.x;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, nwaves = blockDim.x >> 6;
  cplx* A = matptr(c, Ar);
  double* D = vecptr(c, Dr);
  GroupSync g = chain_group(c);

  extern __shared__ double smem[];
  cplx* P = reinterpret_cast<cplx*>(smem);  // [nb][n] panel (rows k0..n-1 used) / staged V
  __shared__ int perm[kQrMaxN], lpos[kQrMaxN];
  __shared__ cplx Tl[kQrMaxNb * kQrMaxNb];  // Tl[j * nb + t] = T(j, t)
  __shared__ cplx taul[kQrMaxNb];
  __shared__ double betal[kQrMaxNb];
  __shared__ double red[8];

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

  for (int k0 = 0, p = 0; k0 < n; k0 += nb, ++p) {
    const int kb = min(nb, n - k0);
    cplx* Tg = c.qr_Tf + (size_t)p * nb * nb;
    if (w == 0) {
      // ---- panel factorisation in LDS
      for (int e = tid; e < kb * (n - k0); e += blockDim.x) {
        const int t = e / (n - k0), r = k0 + (e - t * (n - k0));
        P[(size_t)t * n + r] = A[(size_t)perm[k0 + t] * n + r];
      }
      __syncthreads();
      // per column: wave 0 forms the reflector (zlarfg) alone, one barrier publishes it, every
      // wave applies it to its share of the remaining panel columns, one barrier
      for (int t = 0; t < kb; ++t) {
        const int i = k0 + t;
        cplx* x = P + (size_t)t * n;
        if (wave == 0) {
          double part = 0.0;
          for (int k = i + 1 + lane; k < n; k += 64) part += abs2(x[k]);
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
            for (int k = i + 1 + lane; k < n; k += 64) x[k] = x[k] * sc;
          }
          if (lane == 0) {  // global copies are written after the panel: a store here
            taul[t] = tau;   // would make every later __syncthreads wait for it
            betal[t] = beta;
          }
        }
        __syncthreads();
        // H^H = I - conj(tau) v v^H on the remaining panel columns: wave w takes columns
        // t+1+w, t+5+w, ... together, so their loads and sums overlap
        const cplx ctau = conj(taul[t]);
        {
          constexpr int kC = kQrMaxNb / 4;
          cplx acc[kC];
#pragma unroll
          for (int u = 0; u < kC; ++u) acc[u] = mk(0, 0);
          for (int k = i + lane; k < n; k += 64) {
            const cplx vk = (k == i) ? mk(1, 0) : x[k];
#pragma unroll
            for (int u = 0; u < kC; ++u) {
              const int cc = t + 1 + wave + 4 * u;
              if (cc < kb) cfma_cj(acc[u], vk, P[(size_t)cc * n + k]);
            }
          }
          cplx f[kC];
#pragma unroll
          for (int u = 0; u < kC; ++u) f[u] = (t + 1 + wave + 4 * u < kb) ? ctau * wsum(acc[u]) : mk(0, 0);
          for (int k = i + lane; k < n; k += 64) {
            const cplx vk = (k == i) ? mk(1, 0) : x[k];
#pragma unroll
            for (int u = 0; u < kC; ++u) {
              const int cc = t + 1 + wave + 4 * u;
              if (cc < kb) P[(size_t)cc * n + k] = P[(size_t)cc * n + k] - f[u] * vk;
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
      // first every inner product V(:, j)^H v_t, j < t, into Tl(j, t)
      // (wave w takes t = w, w+4, ...; the t inner products of one t share v_t's loads)
      for (int t = 1 + wave; t < kb; t += nwaves) {
        const cplx* vt = P + (size_t)t * n;
        const int it = k0 + t;
        cplx acc[kQrMaxNb];
#pragma unroll
        for (int j = 0; j < kQrMaxNb; ++j) acc[j] = mk(0, 0);
        for (int k = it + lane; k < n; k += 64) {
          const cplx b = (k == it) ? mk(1, 0) : vt[k];
#pragma unroll
          for (int j = 0; j < kQrMaxNb; ++j)
            if (j < t) cfma_cj(acc[j], P[(size_t)j * n + k], b);  // rows >= it: below v_j's diagonal
        }
#pragma unroll
        for (int j = 0; j < kQrMaxNb; ++j)
          if (j < t) {
            const cplx sj = wsum(acc[j]);
            if (lane == 0) Tl[j * nb + t] = sj;
          }
      }
      __syncthreads();
      // column t of T: thread j < t forms sum_{l=j}^{t-1} T(j, l) ip(l, t) from the inner products
      // still held in rows l >= j of column t, then overwrites row j; columns in order
      for (int t = 0; t < kb; ++t) {
        if (tid < t) {
          cplx acc = mk(0, 0);
          for (int l = tid; l < t; ++l) cfma(acc, Tl[tid * nb + l], Tl[l * nb + t]);
          Tl[tid * nb + t] = -(taul[t] * acc);
        } else if (tid == t) {
          Tl[t * nb + t] = taul[t];
        } else if (tid < kb) {
          Tl[tid * nb + t] = mk(0, 0);
        }
        __syncthreads();
      }
      for (int e = tid; e < kb * kb; e += blockDim.x) Tg[e] = Tl[e];
      for (int e = tid; e < kb * (n - k0); e += blockDim.x) {
        const int t = e / (n - k0), r = k0 + (e - t * (n - k0));
        A[(size_t)perm[k0 + t] * n + r] = P[(size_t)t * n + r];
      }
    }
    group_sync(g);
    if (group_failed(g)) return;
    if (k0 + kb >= n) break;
    // ---- trailing update: a <- (I - V T^H V^H) a for the own columns of logical position >= k0 + kb
    for (int e = tid; e < kb * (n - k0); e += blockDim.x) {
      const int t = e / (n - k0), r = k0 + (e - t * (n - k0));
      const cplx val = A[(size_t)perm[k0 + t] * n + r];
      P[(size_t)t * n + r] = (r < k0 + t) ? mk(0, 0) : (r == k0 + t) ? mk(1, 0) : val;
    }
    for (int e = tid; e < kb * kb; e += blockDim.x) Tl[e] = Tg[e];
    __syncthreads();
    for (int col = w + W * wave; col < n; col += W * nwaves) {
      if (lpos[col] < k0 + kb) continue;
      cplx* a = A + (size_t)col * n;
      // the column's rows k0.. in registers (all loads in flight together)
      cplx av[kQrMaxN / 64];
#pragma unroll
      for (int q = 0; q < kQrMaxN / 64; ++q) {
        const int k = k0 + lane + 64 * q;
        av[q] = k < n ? a[k] : mk(0, 0);
      }
      cplx wv[kQrMaxNb];
#pragma unroll
      for (int j = 0; j < kQrMaxNb; ++j) wv[j] = mk(0, 0);
#pragma unroll
      for (int q = 0; q < kQrMaxN / 64; ++q) {
        const int k = k0 + lane + 64 * q;
        if (k < n) {
#pragma unroll
          for (int j = 0; j < kQrMaxNb; ++j)
            if (j < kb) cfma_cj(wv[j], P[(size_t)j * n + k], av[q]);
        }
      }
#pragma unroll
      for (int j = 0; j < kQrMaxNb; ++j)
        if (j < kb) wv[j] = wsum(wv[j]);
      // z = T^H w (T upper triangular)
      cplx z[kQrMaxNb];
#pragma unroll
      for (int j = 0; j < kQrMaxNb; ++j) {
        cplx acc = mk(0, 0);
#pragma unroll
        for (int l = 0; l < kQrMaxNb; ++l)
          if (l <= j && j < kb) cfma_cj(acc, Tl[l * nb + j], wv[l]);
        z[j] = acc;
      }
#pragma unroll
      for (int q = 0; q < kQrMaxN / 64; ++q) {
        const int k = k0 + lane + 64 * q;
        if (k < n) {
          cplx ak = av[q];
#pragma unroll
          for (int j = 0; j < kQrMaxNb; ++j)
            if (j < kb) ak = ak - P[(size_t)j * n + k] * z[j];
          a[k] = ak;
        }
      }
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
// workgroup owns `cb` columns of B in LDS, applies the panels' Q_p^H in order, then
// back-substitutes with R; no synchronisation between workgroups.
__global__ __launch_bounds__(256) void k_qr_solve(const ChainDesc* chains, MatRef Ar, MatRef Br, int nb, int cb) {
  const ChainDesc& c = chains[blockIdx.y];
  const int n = c.n;
  const int j0 = blockIdx.x * cb;
  if (j0 >= n) return;
  const int ncol = min(cb, n - j0);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, nwaves = blockDim.x >> 6;
  const cplx* A = matptr(c, Ar);
  cplx* B = matptr(c, Br);
  extern __shared__ double smem[];
  cplx* bc = reinterpret_cast<cplx*>(smem);  // [cb][n]
  cplx* V = bc + (size_t)cb * n;              // [nb][n]
  __shared__ cplx Tl[kQrMaxNb * kQrMaxNb];
  for (int e = tid; e < ncol * n; e += blockDim.x) bc[e] = B[(size_t)j0 * n + e];
  for (int k0 = 0, p = 0; k0 < n; k0 += nb, ++p) {
    const int kb = min(nb, n - k0);
    __syncthreads();
    for (int e = tid; e < kb * (n - k0); e += blockDim.x) {
      const int t = e / (n - k0), r = k0 + (e - t * (n - k0));
      V[(size_t)t * n + r] = (r == k0 + t) ? mk(1, 0) : (r < k0 + t) ? mk(0, 0) : A[(size_t)(k0 + t) * n + r];
    }
    for (int e = tid; e < kb * kb; e += blockDim.x) Tl[e] = c.qr_Tf[(size_t)p * nb * nb + e];
    __syncthreads();
    for (int jl = wave; jl < ncol; jl += nwaves) {
      cplx* a = bc + (size_t)jl * n;
      cplx wv[kQrMaxNb];
#pragma unroll
      for (int j = 0; j < kQrMaxNb; ++j) wv[j] = mk(0, 0);
      for (int k = k0 + lane; k < n; k += 64) {
        const cplx ak = a[k];
#pragma unroll
        for (int j = 0; j < kQrMaxNb; ++j)
          if (j < kb) cfma_cj(wv[j], V[(size_t)j * n + k], ak);
      }
#pragma unroll
      for (int j = 0; j < kQrMaxNb; ++j)
        if (j < kb) wv[j] = wsum(wv[j]);
      cplx z[kQrMaxNb];
#pragma unroll
      for (int j = 0; j < kQrMaxNb; ++j) {
        cplx acc = mk(0, 0);
#pragma unroll
        for (int l = 0; l < kQrMaxNb; ++l)
          if (l <= j && j < kb) cfma_cj(acc, Tl[l * nb + j], wv[l]);
        z[j] = acc;
      }
      for (int k = k0 + lane; k < n; k += 64) {
        cplx ak = a[k];
#pragma unroll
        for (int j = 0; j < kQrMaxNb; ++j)
          if (j < kb) ak = ak - V[(size_t)j * n + k] * z[j];
        a[k] = ak;
      }
    }
  }
  __syncthreads();
  // back substitution R y = z in blocks of 16 rows from the bottom: one wave per column solves
  // the 16 x 16 diagonal block (lane r holds row lo + r), then all threads subtract the block's
  // contribution from the rows above, reading R's block column once, coalesced
  for (int hi = n; hi > 0; hi -= 16) {
    const int lo = max(0, hi - 16), bs = hi - lo;
    __syncthreads();
    // the 16 x 16 diagonal block of R and its reciprocal diagonal, staged once
    for (int e = tid; e < bs * bs; e += blockDim.x) {
      const int kk = e / bs, rr = e - kk * bs;
      Tl[kk * 16 + rr] = A[(size_t)(lo + kk) * n + lo + rr];
    }
    __syncthreads();
    for (int jl = wave; jl < ncol; jl += nwaves) {
      cplx* y = bc + (size_t)jl * n;
      const int r = lo + (lane & 15);
      cplx zr = (lane < bs) ? y[r] : mk(0, 0);
      for (int k = bs - 1; k >= 0; --k) {
        const cplx yk = shfl_c(zr, k) * cinv(Tl[k * 16 + k]);
        if (lane < k) zr = zr - Tl[k * 16 + (lane & 15)] * yk;
        if (lane == k) zr = yk;
      }
      if (lane < bs) y[r] = zr;
    }
    __syncthreads();
    for (int e = tid; e < lo * ncol; e += blockDim.x) {
      const int jl = e / lo, r = e - jl * lo;
      cplx* y = bc + (size_t)jl * n;
      cplx acc = y[r];
      for (int k = lo; k < hi; ++k) acc = acc - A[(size_t)k * n + r] * y[k];
      y[r] = acc;
    }
  }
  __syncthreads();
  for (int e = tid; e < ncol * n; e += blockDim.x) B[(size_t)j0 * n + e] = bc[e];
}

int qr_panel_width(int nmax) { return nmax <= 512 ? 16 : 8; }

hipError_t launch_qr_blocked(hipStream_t s, const ChainDesc* chains, int nchains, int nmax, MatRef A, VecRef D,
                             int prepivot, int wg) {
  if (nmax > kQrMaxN) return hipErrorInvalidValue;
  int nb = qr_panel_width(nmax);
  const size_t lds = (size_t)nb * nmax * sizeof(cplx);
  static size_t attr = 0;
  if (lds > 65536 && lds > attr) {
    hipError_t e =
        hipFuncSetAttribute((const void*)k_qr_blocked, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    attr = lds;
  }
  void* args[] = {(void*)&chains, (void*)&A, (void*)&D, (void*)&prepivot, (void*)&nb};
  hipError_t e = hipLaunchCooperativeKernel((const void*)k_qr_blocked, dim3(wg, nchains), dim3(256), args, lds, s);
  if (e != hipSuccess) (void)hipGetLastError();
  return e;
}

hipError_t launch_qr_solve(hipStream_t s, const ChainDesc* chains, int nchains, int nmax, MatRef A, MatRef B) {
  if (nmax > kQrMaxN) return hipErrorInvalidValue;
  const int nb = qr_panel_width(nmax);
  const int cb = nmax <= 512 ? 4 : 1;
  const size_t lds = (size_t)(cb + nb) * nmax * sizeof(cplx);
  static size_t attr = 0;
  if (lds > 65536 && lds > attr) {
    hipError_t e = hipFuncSetAttribute((const void*)k_qr_solve, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    attr = lds;
  }
  hipLaunchKernelGGL(k_qr_solve, dim3((nmax + cb - 1) / cb, nchains), dim3(256), lds, s, chains, A, B, nb, cb);
  return hipGetLastError();
}

}  // namespace dqmc
