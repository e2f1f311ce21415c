// UDT decomposition A = U diag(D) T by column-pivoted Householder QR, and the
// in-place inverse by Gauss-Jordan elimination with partial pivoting.
//
// Reference: decompose_udt! (linalg.jl:20-39) = LAPACK zgeqp3 + zungqr with
// D = |Re R_ii| and T = D^-1 R P^T; inv (Julia getrf + getri).
//
// Pivoted QR on a chain group of W workgroups (cooperative launch):
//   * columns never move in memory; a logical->physical permutation lives in
//     every workgroup's LDS and pivoting swaps its entries;
//   * physical column c belongs to workgroup c % W for the whole factorization;
//   * at the end of each step every workgroup publishes its best remaining
//     column (norm, logical position, values) to a double-buffered slot, so the
//     next step's pivot choice and Householder vector are computed redundantly by
//     every workgroup from the published copy: one group barrier per column;
//   * pivot rule and partial-norm downdating follow LAPACK zlaqp2 (first maximum,
//     tol3z = sqrt(eps) recomputation threshold).
// Q is then formed column-independently (Q e_j = H_0 ... H_j e_j) with blocks of
// reflectors staged in LDS; T is an elementwise rescale of R.
#include "chain.hpp"

namespace dqmc {

constexpr int kMaxN = 1024;

__device__ inline double cabs1(cplx a) { return fabs(a.re) + fabs(a.im); }

// Wave64 dot conj(v) . x over rows [r0, n) of a column, x strided by 1.
__device__ inline cplx wave_dotc(const cplx* v, const cplx* x, int r0, int n) {
  const int lane = threadIdx.x & 63;
  cplx acc = mk(0, 0);
  for (int k = r0 + lane; k < n; k += 64) cfma_cj(acc, v[k], x[k]);
  return wave_sum(acc);
}
__device__ inline double wave_norm2(const cplx* x, int r0, int n) {
  const int lane = threadIdx.x & 63;
  double acc = 0.0;
  for (int k = r0 + lane; k < n; k += 64) acc += abs2(x[k]);
  return wave_sum(acc);
}

__global__ __launch_bounds__(256) void k_qrp(const ChainDesc* chains, MatRef Ar, VecRef Dr) {
  const ChainDesc& c = chains[blockIdx.y];
  const int n = c.n;
  const int W = gridDim.x, w = blockIdx.x;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, nwaves = blockDim.x >> 6;
  cplx* A = matptr(c, Ar);
  double* D = vecptr(c, Dr);
  GroupSync g = chain_group(c);
  // partial column norms of the own columns, indexed by physical column (LDS)
  __shared__ double vn1[kMaxN], vn2[kMaxN];
  const double tol3z = sqrt(0.5 * 2.220446049250313e-16);

  __shared__ int perm[kMaxN], lpos[kMaxN];
  __shared__ cplx v[kMaxN];
  __shared__ double red[8];
  __shared__ int s_pos, s_win;
  __shared__ double s_best[4];
  __shared__ int s_bestc[4], s_bestl[4];

  for (int i = tid; i < n; i += blockDim.x) perm[i] = lpos[i] = i;
  // initial column norms of own columns
  for (int col = w + W * wave; col < n; col += W * nwaves) {
    const double nr = sqrt(wave_norm2(A + (size_t)col * n, 0, n));
    if (lane == 0) {
      vn1[col] = nr;
      vn2[col] = nr;
    }
  }
  __syncthreads();

  // publish best own column among logical positions >= first
  auto publish = [&](int first, int buf) {
    double best = -1.0;
    int bestc = -1, bestl = n;
    if (lane == 0) {
      for (int col = w + W * wave; col < n; col += W * nwaves) {
        const int lp = lpos[col];
        if (lp < first) continue;
        const double nv = vn1[col];
        if (nv > best || (nv == best && lp < bestl)) {
          best = nv;
          bestc = col;
          bestl = lp;
        }
      }
      s_best[wave] = best;
      s_bestc[wave] = bestc;
      s_bestl[wave] = bestl;
    }
    __syncthreads();
    if (tid == 0) {
      int bw = 0;
      for (int q = 1; q < nwaves; ++q) {
        if (s_best[q] > s_best[bw] || (s_best[q] == s_best[bw] && s_bestl[q] < s_bestl[bw])) bw = q;
      }
      c.qr_cand_norm[buf * W + w] = s_bestc[bw] >= 0 ? s_best[bw] : -1.0;
      c.qr_cand_pos[buf * W + w] = s_bestc[bw] >= 0 ? s_bestl[bw] : -1;
      s_win = s_bestc[bw];
    }
    __syncthreads();
    const int bc = s_win;
    if (bc >= 0) {
      cplx* dst = c.qr_cand + ((size_t)buf * W + w) * n;
      const cplx* src = A + (size_t)bc * n;
      for (int k = first + tid; k < n; k += blockDim.x) dst[k] = src[k];
    }
  };

  publish(0, 0);
  group_sync(g);

  for (int i = 0; i < n; ++i) {
    if (group_failed(g)) return;
    const int buf = i & 1;
    // ---- pivot: first maximum of the partial norms over logical positions >= i.
    // Every candidate slot is read by its own thread (one memory round trip), then
    // reduced in wave 0 by (norm descending, logical position ascending).
    if (wave == 0) {
      double bn = -1.0;
      int bp = n, bw = -1;
      for (int q = lane; q < W; q += 64) {
        const int pq = c.qr_cand_pos[buf * W + q];
        const double nq = c.qr_cand_norm[buf * W + q];
        if (pq >= 0 && (bw < 0 || nq > bn || (nq == bn && pq < bp))) {
          bn = nq;
          bp = pq;
          bw = q;
        }
      }
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) {
        const double on = __shfl_xor(bn, off, 64);
        const int op = __shfl_xor(bp, off, 64);
        const int ow = __shfl_xor(bw, off, 64);
        if (ow >= 0 && (bw < 0 || on > bn || (on == bn && op < bp))) {
          bn = on;
          bp = op;
          bw = ow;
        }
      }
      if (lane == 0) {
        s_win = bw;
        s_pos = bp;
      }
    }
    __syncthreads();
    if (tid == 0) {
      const int bp = s_pos;
      const int pc = perm[bp], pi = perm[i];
      perm[i] = pc;
      perm[bp] = pi;
      lpos[pc] = i;
      lpos[pi] = bp;
    }
    __syncthreads();
    const int pc = perm[i];
    const cplx* x = c.qr_cand + ((size_t)buf * W + s_win) * n;
    // ---- Householder reflector (zlarfg) from rows i..n-1 of the pivot column
    double part = 0.0;
    for (int k = i + 1 + tid; k < n; k += blockDim.x) {
      const cplx xv = x[k];
      v[k] = xv;
      part += abs2(xv);
    }
    const double xnorm = sqrt(block_sum(part, red));
    const cplx alpha = x[i];
    cplx tau;
    double beta;
    if (xnorm == 0.0 && alpha.im == 0.0) {
      tau = mk(0, 0);
      beta = alpha.re;
    } else {
      beta = -copysign(sqrt(alpha.re * alpha.re + alpha.im * alpha.im + xnorm * xnorm), alpha.re);
      tau = mk((beta - alpha.re) / beta, -alpha.im / beta);
      const cplx sc = cinv(alpha - mk(beta, 0));
      for (int k = i + 1 + tid; k < n; k += blockDim.x) v[k] = v[k] * sc;
    }
    if (tid == 0) v[i] = mk(1, 0);
    __syncthreads();
    if (w == 0 && tid == 0) {
      D[i] = fabs(beta);
      c.qr_tau[i] = tau;
    }
    if (pc % W == w) {
      cplx* colp = A + (size_t)pc * n;
      for (int k = i + tid; k < n; k += blockDim.x) colp[k] = (k == i) ? mk(beta, 0) : v[k];
    }
    // ---- apply H^H = I - conj(tau) v v^H to own trailing columns, downdate norms
    const cplx ctau = conj(tau);
    for (int col = w + W * wave; col < n; col += W * nwaves) {
      if (lpos[col] <= i) continue;
      cplx* colp = A + (size_t)col * n;
      const cplx d = wave_dotc(v, colp, i, n);
      const cplx f = ctau * d;
      // update rows i..n-1; keep the new row-i entry (lane 0) and the tail norm in registers
      cplx first = mk(0, 0);
      double tail = 0.0;
      for (int k = i + lane; k < n; k += 64) {
        const cplx nv = colp[k] - f * v[k];
        colp[k] = nv;
        if (k == i)
          first = nv;
        else
          tail += abs2(nv);
      }
      const double aic = sqrt(abs2(mk(__shfl(first.re, 0, 64), __shfl(first.im, 0, 64))));
      const double tailnorm = sqrt(wave_sum(tail));
      const double v1 = vn1[col];
      if (v1 != 0.0 && lane == 0) {
        double temp = 1.0 - (aic / v1) * (aic / v1);
        temp = fmax(temp, 0.0);
        const double ratio = v1 / vn2[col];
        const double temp2 = temp * ratio * ratio;
        if (temp2 <= tol3z) {
          // partial norm lost its accuracy: recompute from the trailing rows (zlaqp2)
          const double nr = (i < n - 1) ? tailnorm : 0.0;
          vn1[col] = nr;
          vn2[col] = nr;
        } else {
          vn1[col] = v1 * sqrt(temp);
        }
      }
    }
    __syncthreads();
    if (i + 1 < n) publish(i + 1, buf ^ 1);
    group_sync(g);
  }
  if (w == 0) {
    for (int k = tid; k < n; k += blockDim.x) {
      c.qr_perm[k] = perm[k];
      c.qr_lpos[k] = lpos[k];
    }
  }
}

// T = D^-1 R P^T : T[r, col] = R[r, col] / D[r] for r <= lpos[col], else 0.
__global__ __launch_bounds__(256) void k_form_T(const ChainDesc* chains, MatRef Ar, VecRef Dr, MatRef Tr) {
  const ChainDesc& c = chains[blockIdx.y];
  const int n = c.n;
  const size_t nn = (size_t)n * n;
  const cplx* A = matptr(c, Ar);
  const double* D = vecptr(c, Dr);
  cplx* T = matptr(c, Tr);
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < nn; e += (size_t)gridDim.x * blockDim.x) {
    const int col = (int)(e / n), r = (int)(e - (size_t)col * n);
    T[e] = (r <= c.qr_lpos[col]) ? (1.0 / D[r]) * A[e] : mk(0, 0);
  }
}

// Q = H_0 H_1 ... H_{n-1}: each workgroup owns CB columns in LDS and streams
// blocks of RB reflectors through LDS from the last to the first.
__global__ __launch_bounds__(256) void k_form_Q(const ChainDesc* chains, MatRef Ar, MatRef Qr, int cb) {
  const ChainDesc& c = chains[blockIdx.y];
  const int n = c.n;
  const int j0 = blockIdx.x * cb;
  if (j0 >= n) return;
  const int ncol = min(cb, n - j0);
  const int rb = cb;
  const cplx* A = matptr(c, Ar);
  cplx* Q = matptr(c, Qr);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, nwaves = blockDim.x >> 6;
  extern __shared__ double smem[];
  cplx* qc = reinterpret_cast<cplx*>(smem);  // [cb][n]
  cplx* vr = qc + (size_t)cb * n;             // [rb][n]
  __shared__ cplx taus[16];

  for (int e = tid; e < ncol * n; e += blockDim.x) {
    const int jl = e / n, r = e - jl * n;
    qc[e] = mk(r == j0 + jl ? 1.0 : 0.0, 0.0);
  }
  const int jmax = j0 + ncol - 1;
  // reflectors i > jmax leave every owned column unchanged
  for (int ihi = jmax; ihi >= 0; ihi -= rb) {
    const int ilo = max(0, ihi - rb + 1);
    __syncthreads();
    for (int e = tid; e < (ihi - ilo + 1) * n; e += blockDim.x) {
      const int t = e / n, r = e - t * n;
      const int i = ilo + t;
      const int pcol = c.qr_perm[i];
      cplx val;
      if (r < i)
        val = mk(0, 0);
      else if (r == i)
        val = mk(1, 0);
      else
        val = A[(size_t)pcol * n + r];
      vr[(size_t)t * n + r] = val;
    }
    if (tid < ihi - ilo + 1) taus[tid] = c.qr_tau[ilo + tid];
    __syncthreads();
    for (int jl = wave; jl < ncol; jl += nwaves) {
      const int j = j0 + jl;
      cplx* q = qc + (size_t)jl * n;
      for (int i = min(ihi, j); i >= ilo; --i) {
        const cplx* vi = vr + (size_t)(i - ilo) * n;
        const cplx d = wave_dotc(vi, q, i, n);
        const cplx f = taus[i - ilo] * d;
        for (int k = i + lane; k < n; k += 64) q[k] = q[k] - f * vi[k];
        __builtin_amdgcn_wave_barrier();
      }
    }
  }
  __syncthreads();
  for (int e = tid; e < ncol * n; e += blockDim.x) Q[(size_t)j0 * n + e] = qc[e];
}

static hipError_t coop_launch(const void* fn, dim3 grid, void** args, hipStream_t s) {
  return hipLaunchCooperativeKernel(fn, grid, dim3(256), args, 0, s);
}

hipError_t launch_qrp(hipStream_t s, const ChainDesc* chains, int nchains, int nmax, MatRef A, VecRef D, int wg) {
  if (nmax > kMaxN) return hipErrorInvalidValue;
  void* args[] = {(void*)&chains, (void*)&A, (void*)&D};
  hipError_t e = coop_launch((const void*)k_qrp, dim3(wg, nchains), args, s);
  if (e != hipSuccess) (void)hipGetLastError();
  return e;
}

hipError_t launch_udt(hipStream_t s, const ChainDesc* chains, int nchains, int nmax, MatRef A, VecRef D,
                      MatRef Q, MatRef T, int wg) {
  hipError_t e = launch_qrp(s, chains, nchains, nmax, A, D, wg);
  if (e != hipSuccess) return e;
  return launch_form_TQ(s, chains, nchains, nmax, A, D, Q, T);
}

hipError_t launch_form_T(hipStream_t s, const ChainDesc* chains, int nchains, int nmax, MatRef A, VecRef D,
                         MatRef T) {
  if (nmax > kMaxN) return hipErrorInvalidValue;
  const size_t nn = (size_t)nmax * nmax;
  int blocks = (int)((nn + 255) / 256);
  if (blocks > 1024) blocks = 1024;
  hipLaunchKernelGGL(k_form_T, dim3(blocks, nchains), dim3(256), 0, s, chains, A, D, T);
  return hipGetLastError();
}

hipError_t launch_form_TQ(hipStream_t s, const ChainDesc* chains, int nchains, int nmax, MatRef A, VecRef D,
                          MatRef Q, MatRef T) {
  hipError_t e0 = launch_form_T(s, chains, nchains, nmax, A, D, T);
  if (e0 != hipSuccess) return e0;
  const int cb = nmax <= 512 ? 8 : 4;
  const size_t lds = (size_t)2 * cb * nmax * sizeof(cplx);
  {
    hipError_t e = ensure_dynamic_lds((const void*)k_form_Q, lds);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(k_form_Q, dim3((nmax + cb - 1) / cb, nchains), dim3(256), lds, s, chains, A, Q, cb);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Gauss-Jordan inverse with partial pivoting (pivot = first max of |re|+|im|,
// as izamax in zgetf2), one group barrier per column.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_gj_inverse(const ChainDesc* chains, MatRef Xr) {
  const ChainDesc& c = chains[blockIdx.y];
  const int n = c.n;
  const int W = gridDim.x, w = blockIdx.x;
  const int tid = threadIdx.x;
  cplx* X = matptr(c, Xr);
  GroupSync g = chain_group(c);
  __shared__ cplx col[kMaxN];
  __shared__ cplx tv[kMaxN / 2];
  __shared__ double rv[256];
  __shared__ int ri[256];
  __shared__ int s_p;
  __shared__ double s_logdet;
  __shared__ int pivl[kMaxN], idxl[kMaxN];
  const int nown = (n - w + W - 1) / W;  // own columns: w, w+W, ...

  if (tid == 0) s_logdet = 0.0;
  if (w == 0)
    for (int r = tid; r < n; r += blockDim.x) c.inv_bcast[r] = X[r];
  group_sync(g);

  for (int k = 0; k < n; ++k) {
    if (group_failed(g)) return;
    const int buf = k & 1;
    const cplx* bc = c.inv_bcast + (size_t)buf * n;
    for (int r = tid; r < n; r += blockDim.x) col[r] = bc[r];
    __syncthreads();
    // pivot search over rows k..n-1: first maximum
    double bv = -1.0;
    int bi = n;
    for (int r = k + tid; r < n; r += blockDim.x) {
      const double a = cabs1(col[r]);
      if (a > bv) {
        bv = a;
        bi = r;
      }
    }
    rv[tid] = bv;
    ri[tid] = bi;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
      if (tid < off) {
        const double ov = rv[tid + off];
        const int oi = ri[tid + off];
        if (ov > rv[tid] || (ov == rv[tid] && oi < ri[tid])) {
          rv[tid] = ov;
          ri[tid] = oi;
        }
      }
      __syncthreads();
    }
    if (tid == 0) {
      s_p = ri[0];
      pivl[k] = s_p;
    }
    __syncthreads();
    const int p = s_p;
    if (p != k) {
      // swap rows k and p in own columns and in the broadcast column copy
      for (int jl = tid; jl < nown; jl += blockDim.x) {
        const int j = w + jl * W;
        const cplx a = X[(size_t)j * n + k];
        X[(size_t)j * n + k] = X[(size_t)j * n + p];
        X[(size_t)j * n + p] = a;
      }
      __syncthreads();
      if (tid == 0) {
        const cplx a = col[k];
        col[k] = col[p];
        col[p] = a;
      }
      __syncthreads();
    }
    const cplx piv = col[k];
    const cplx pinv = cinv(piv);
    if (tid == 0) s_logdet += 0.5 * log(abs2(piv));
    // multipliers of the pivot row
    for (int jl = tid; jl < nown; jl += blockDim.x) {
      const int j = w + jl * W;
      tv[jl] = (j == k) ? mk(0, 0) : X[(size_t)j * n + k] * pinv;
    }
    __syncthreads();
    for (int e = tid; e < nown * n; e += blockDim.x) {
      const int jl = e / n, r = e - jl * n;
      const int j = w + jl * W;
      cplx* xp = X + (size_t)j * n + r;
      if (j == k) {
        *xp = (r == k) ? pinv : -(col[r] * pinv);
      } else if (r == k) {
        *xp = tv[jl];
      } else {
        *xp = *xp - col[r] * tv[jl];
      }
    }
    __syncthreads();
    if (k + 1 < n && ((k + 1) % W) == w) {
      cplx* dst = c.inv_bcast + (size_t)(buf ^ 1) * n;
      const cplx* src = X + (size_t)(k + 1) * n;
      for (int r = tid; r < n; r += blockDim.x) dst[r] = src[r];
    }
    group_sync(g);
  }
  if (w == 0 && tid == 0) c.scal[S_LOGDET_A] = s_logdet;
  // Undo the row interchanges as column interchanges (last first). The composed
  // permutation is built in LDS; every workgroup gathers its own output columns
  // into M_INVT (all columns are visible after the last group barrier).
  if (tid == 0) {
    for (int j = 0; j < n; ++j) idxl[j] = j;
    for (int k = n - 1; k >= 0; --k) {
      const int p = pivl[k];
      const int t = idxl[k];
      idxl[k] = idxl[p];
      idxl[p] = t;
    }
  }
  __syncthreads();
  cplx* Y = c.mat[M_INVT];
  for (int e = tid; e < nown * n; e += blockDim.x) {
    const int jl = e / n, r = e - jl * n;
    const int j = w + jl * W;
    Y[(size_t)j * n + r] = X[(size_t)idxl[j] * n + r];
  }
}

hipError_t launch_inverse_gather(hipStream_t s, const ChainDesc* chains, int nchains, int nmax, MatRef X, int wg) {
  if (nmax > kMaxN || wg < 1) return hipErrorInvalidValue;
  if ((nmax + wg - 1) / wg > kMaxN / 2) return hipErrorInvalidValue;
  void* args[] = {(void*)&chains, (void*)&X};
  hipError_t e = coop_launch((const void*)k_gj_inverse, dim3(wg, nchains), args, s);
  if (e != hipSuccess) (void)hipGetLastError();
  return e;
}

hipError_t launch_inverse(hipStream_t s, const ChainDesc* chains, int nchains, int nmax, MatRef X, int wg) {
  hipError_t e = launch_inverse_gather(s, chains, nchains, nmax, X, wg);
  if (e != hipSuccess) return e;
  return launch_copy(s, chains, nchains, nmax, X, MatRef{MK_NAMED, M_INVT});
}

}  // namespace dqmc
