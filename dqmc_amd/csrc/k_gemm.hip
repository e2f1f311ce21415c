// Complex FP64 GEMM on the FP64 matrix cores, plus the elementwise helpers the
// stack algebra needs (copies, diagonal scalings, adjoint, max-abs difference).
//
// Reference call sites: the dense products of stack.jl:291,311,346-366 and
// linalg.jl:237-283 (zgemm through LinearAlgebra.mul!), with the Diagonal
// rmul!/lmul! scalings folded into the GEMM epilogue.
//
// Tiling: a 256-thread workgroup owns a 32x32 complex tile of C; each wave64 owns
// a 16x16 complex sub-tile accumulated with v_mfma_f64_16x16x4 (four real MFMAs
// per complex k-step: re*re, -im*im, re*im, im*re). K advances in chunks of 16
// staged through LDS.
#include <stdlib.h>

#include "chain.hpp"

namespace dqmc {

typedef double v4d __attribute__((ext_vector_type(4)));

constexpr int GT = 32;  // C tile edge
constexpr int KT = 16;  // K chunk

// Row of accumulator element v held by `lane` in v_mfma_f64_16x16x4 (column = lane % 16):
// rows are interleaved, lane group g = lane / 16 holds rows g, g + 4, g + 8, g + 12
// (measured on gfx950 with an in-bounds probe of the builtin).
__device__ inline int kMfmaRow(int lane, int v) { return (lane >> 4) + 4 * v; }

// Shared tile loader: As[r][kk] = op(A)[i0 + r][k0 + kk], Bs[kk][c] = op(B)[k0 + kk][j0 + c].
template <int OPA, int OPB>
__device__ inline void load_tiles(const cplx* A, const cplx* B, int n, int i0, int j0, int k0,
                                  double (*As_re)[KT + 1], double (*As_im)[KT + 1], double (*Bs_re)[GT + 1],
                                  double (*Bs_im)[GT + 1]) {
  const int tid = threadIdx.x;
#pragma unroll
  for (int t = 0; t < 2; ++t) {
    const int e = tid + 256 * t;
    int r, kk;
    if (OPA == OP_N) {
      r = e & 31;
      kk = e >> 5;
    } else {
      kk = e & 15;
      r = e >> 4;
    }
    const int gi = i0 + r, gk = k0 + kk;
    cplx v = mk(0, 0);
    if (gi < n && gk < n) v = (OPA == OP_N) ? A[(size_t)gk * n + gi] : conj(A[(size_t)gi * n + gk]);
    As_re[r][kk] = v.re;
    As_im[r][kk] = v.im;
  }
#pragma unroll
  for (int t = 0; t < 2; ++t) {
    const int e = tid + 256 * t;
    int kk, cc;
    if (OPB == OP_N) {
      kk = e & 15;
      cc = e >> 4;
    } else {
      cc = e & 31;
      kk = e >> 5;
    }
    const int gk = k0 + kk, gj = j0 + cc;
    cplx v = mk(0, 0);
    if (gk < n && gj < n) v = (OPB == OP_N) ? B[(size_t)gj * n + gk] : conj(B[(size_t)gk * n + gj]);
    Bs_re[kk][cc] = v.re;
    Bs_im[kk][cc] = v.im;
  }
}

template <int OPA, int OPB>
__global__ __launch_bounds__(256) void k_gemm(const ChainDesc* chains, MatRef Cr, MatRef Ar, MatRef Br, VecRef dlr,
                                              VecRef drr, double beta) {
  const ChainDesc& c = chains[blockIdx.z];
  const int n = c.n;
  const int i0 = blockIdx.y * GT, j0 = blockIdx.x * GT;
  if (i0 >= n || j0 >= n) return;
  const cplx* A = matptr(c, Ar);
  const cplx* B = matptr(c, Br);
  cplx* C = matptr(c, Cr);
  const double* dl = dlr.kind >= 0 ? vecptr(c, dlr) : nullptr;
  const double* dr = drr.kind >= 0 ? vecptr(c, drr) : nullptr;

  __shared__ double As_re[GT][KT + 1], As_im[GT][KT + 1];
  __shared__ double Bs_re[KT][GT + 1], Bs_im[KT][GT + 1];

  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int wr = w >> 1, wc = w & 1;
  v4d acc_re = {0, 0, 0, 0}, acc_im = {0, 0, 0, 0};

  for (int k0 = 0; k0 < n; k0 += KT) {
    // A tile: rows i0..i0+31, k k0..k0+15
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      const int e = tid + 256 * t;
      int r, kk;
      if (OPA == OP_N) {
        r = e & 31;
        kk = e >> 5;
      } else {
        kk = e & 15;
        r = e >> 4;
      }
      const int gi = i0 + r, gk = k0 + kk;
      cplx v = mk(0, 0);
      if (gi < n && gk < n) {
        if (OPA == OP_N)
          v = A[(size_t)gk * n + gi];
        else
          v = conj(A[(size_t)gi * n + gk]);
      }
      As_re[r][kk] = v.re;
      As_im[r][kk] = v.im;
    }
    // B tile: k k0..k0+15, cols j0..j0+31
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      const int e = tid + 256 * t;
      int kk, cc;
      if (OPB == OP_N) {
        kk = e & 15;
        cc = e >> 4;
      } else {
        cc = e & 31;
        kk = e >> 5;
      }
      const int gk = k0 + kk, gj = j0 + cc;
      cplx v = mk(0, 0);
      if (gk < n && gj < n) {
        if (OPB == OP_N)
          v = B[(size_t)gj * n + gk];
        else
          v = conj(B[(size_t)gk * n + gj]);
      }
      Bs_re[kk][cc] = v.re;
      Bs_im[kk][cc] = v.im;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < KT; kk += 4) {
      const int ar = wr * 16 + (lane & 15), ak = kk + (lane >> 4);
      const int bk = kk + (lane >> 4), bc = wc * 16 + (lane & 15);
      const double a_re = As_re[ar][ak], a_im = As_im[ar][ak];
      const double b_re = Bs_re[bk][bc], b_im = Bs_im[bk][bc];
      acc_re = __builtin_amdgcn_mfma_f64_16x16x4f64(a_re, b_re, acc_re, 0, 0, 0);
      acc_re = __builtin_amdgcn_mfma_f64_16x16x4f64(-a_im, b_im, acc_re, 0, 0, 0);
      acc_im = __builtin_amdgcn_mfma_f64_16x16x4f64(a_re, b_im, acc_im, 0, 0, 0);
      acc_im = __builtin_amdgcn_mfma_f64_16x16x4f64(a_im, b_re, acc_im, 0, 0, 0);
    }
    __syncthreads();
  }
  const int col = j0 + wc * 16 + (lane & 15);
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const int row = i0 + wr * 16 + kMfmaRow(lane, v);
    if (row < n && col < n) {
      double s = 1.0;
      if (dl) s *= dl[row];
      if (dr) s *= dr[col];
      cplx val = mk(s * acc_re[v], s * acc_im[v]);
      const size_t o = (size_t)col * n + row;
      if (beta != 0.0) val = val + beta * C[o];
      C[o] = val;
    }
  }
}

hipError_t launch_gemm(hipStream_t s, const ChainDesc* chains, int nchains, int nmax, MatRef C, MatRef A,
                       GemmOp opA, MatRef B, GemmOp opB, VecRef dl, VecRef dr, double beta) {
  dim3 grid((nmax + GT - 1) / GT, (nmax + GT - 1) / GT, nchains);
  if (opA == OP_N && opB == OP_N)
    hipLaunchKernelGGL((k_gemm<OP_N, OP_N>), grid, dim3(256), 0, s, chains, C, A, B, dl, dr, beta);
  else if (opA == OP_N && opB == OP_C)
    hipLaunchKernelGGL((k_gemm<OP_N, OP_C>), grid, dim3(256), 0, s, chains, C, A, B, dl, dr, beta);
  else if (opA == OP_C && opB == OP_N)
    hipLaunchKernelGGL((k_gemm<OP_C, OP_N>), grid, dim3(256), 0, s, chains, C, A, B, dl, dr, beta);
  else
    hipLaunchKernelGGL((k_gemm<OP_C, OP_C>), grid, dim3(256), 0, s, chains, C, A, B, dl, dr, beta);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Elementwise kernels: grid (blocks, nchains), grid-stride over n*n (or n).
// ---------------------------------------------------------------------------
enum EwOp : int {
  EW_COPY = 0,
  EW_IDENT,
  EW_SCALE_COLS,
  EW_ADD_DIAG,
  EW_ADJOINT,
  EW_SCALE_ROWS_COPY,
  EW_ADD_SCALED_COLS,
  EW_NEGATE
};

__global__ __launch_bounds__(256) void k_elementwise(const ChainDesc* chains, int op, MatRef Dr, MatRef Sr, VecRef vr) {
  const ChainDesc& c = chains[blockIdx.y];
  const int n = c.n;
  const size_t nn = (size_t)n * n;
  cplx* D = matptr(c, Dr);
  if (!D) return;  // a merged sweep's plan skips this chain
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  const size_t start = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (op == EW_COPY) {
    const cplx* S = matptr(c, Sr);
    for (size_t e = start; e < nn; e += stride) D[e] = S[e];
  } else if (op == EW_IDENT) {
    for (size_t e = start; e < nn; e += stride) {
      const size_t col = e / n, row = e - col * n;
      D[e] = mk(row == col ? 1.0 : 0.0, 0.0);
    }
  } else if (op == EW_SCALE_COLS) {
    const double* d = vecptr(c, vr);
    for (size_t e = start; e < nn; e += stride) {
      const size_t col = e / n;
      D[e] = d[col] * D[e];
    }
  } else if (op == EW_ADD_DIAG) {
    const double* d = vecptr(c, vr);
    for (size_t e = start; e < (size_t)n; e += stride) D[e * n + e].re += d[e];
  } else if (op == EW_SCALE_ROWS_COPY) {
    const cplx* S = matptr(c, Sr);
    const double* d = vecptr(c, vr);
    for (size_t e = start; e < nn; e += stride) {
      const size_t col = e / n, row = e - col * n;
      D[e] = d[row] * S[e];
    }
  } else if (op == EW_ADD_SCALED_COLS) {
    const cplx* S = matptr(c, Sr);
    const double* d = vecptr(c, vr);
    for (size_t e = start; e < nn; e += stride) {
      const size_t col = e / n;
      D[e] = D[e] + d[col] * S[e];
    }
  } else if (op == EW_NEGATE) {
    for (size_t e = start; e < nn; e += stride) D[e] = -D[e];
  } else if (op == EW_ADJOINT) {
    const cplx* S = matptr(c, Sr);
    for (size_t e = start; e < nn; e += stride) {
      const size_t col = e / n, row = e - col * n;
      D[e] = conj(S[row * n + col]);
    }
  }
}

static hipError_t ew(hipStream_t s, const ChainDesc* chains, int nchains, int nmax, int op, MatRef D, MatRef S,
                     VecRef v) {
  const size_t nn = (size_t)nmax * nmax;
  int blocks = (int)((nn + 255) / 256);
  if (blocks > 1024) blocks = 1024;
  if (op == EW_ADD_DIAG) blocks = (nmax + 255) / 256;
  hipLaunchKernelGGL(k_elementwise, dim3(blocks, nchains), dim3(256), 0, s, chains, op, D, S, v);
  return hipGetLastError();
}

hipError_t launch_add_scaled_cols(hipStream_t s, const ChainDesc* chains, int nchains, int nmax, MatRef X,
                                  MatRef src, VecRef d) {
  return ew(s, chains, nchains, nmax, EW_ADD_SCALED_COLS, X, src, d);
}
hipError_t launch_negate(hipStream_t s, const ChainDesc* chains, int nchains, int nmax, MatRef X) {
  return ew(s, chains, nchains, nmax, EW_NEGATE, X, X, VecRef{0, 0});
}
hipError_t launch_copy(hipStream_t s, const ChainDesc* chains, int nchains, int nmax, MatRef dst, MatRef src) {
  return ew(s, chains, nchains, nmax, EW_COPY, dst, src, VecRef{0, 0});
}
hipError_t launch_identity(hipStream_t s, const ChainDesc* chains, int nchains, int nmax, MatRef dst, VecRef) {
  return ew(s, chains, nchains, nmax, EW_IDENT, dst, dst, VecRef{0, 0});
}
hipError_t launch_scale_cols(hipStream_t s, const ChainDesc* chains, int nchains, int nmax, MatRef X, VecRef d) {
  return ew(s, chains, nchains, nmax, EW_SCALE_COLS, X, X, d);
}
hipError_t launch_add_diag(hipStream_t s, const ChainDesc* chains, int nchains, int nmax, MatRef X, VecRef d) {
  return ew(s, chains, nchains, nmax, EW_ADD_DIAG, X, X, d);
}
hipError_t launch_scale_rows_copy(hipStream_t s, const ChainDesc* chains, int nchains, int nmax, MatRef dst,
                                  MatRef src, VecRef d) {
  return ew(s, chains, nchains, nmax, EW_SCALE_ROWS_COPY, dst, src, d);
}
hipError_t launch_adjoint(hipStream_t s, const ChainDesc* chains, int nchains, int nmax, MatRef dst, MatRef src) {
  return ew(s, chains, nchains, nmax, EW_ADJOINT, dst, src, VecRef{0, 0});
}

enum VecOp : int { VO_COPY = 0, VO_RECIP, VO_ONES };
__global__ __launch_bounds__(256) void k_vec(const ChainDesc* chains, int op, VecRef Dr, VecRef Sr) {
  const ChainDesc& c = chains[blockIdx.y];
  double* d = vecptr(c, Dr);
  const double* sv = vecptr(c, Sr);
  if (!d) return;  // a merged sweep's plan skips this chain
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < c.n; i += gridDim.x * blockDim.x) {
    if (op == VO_COPY)
      d[i] = sv[i];
    else if (op == VO_RECIP)
      d[i] = 1.0 / sv[i];
    else
      d[i] = 1.0;
  }
}
hipError_t launch_copy_vec(hipStream_t s, const ChainDesc* chains, int nchains, int nmax, VecRef dst, VecRef src) {
  hipLaunchKernelGGL(k_vec, dim3((nmax + 255) / 256, nchains), dim3(256), 0, s, chains, (int)VO_COPY, dst, src);
  return hipGetLastError();
}
hipError_t launch_recip_vec(hipStream_t s, const ChainDesc* chains, int nchains, int nmax, VecRef dst, VecRef src) {
  hipLaunchKernelGGL(k_vec, dim3((nmax + 255) / 256, nchains), dim3(256), 0, s, chains, (int)VO_RECIP, dst, src);
  return hipGetLastError();
}
hipError_t launch_ones_vec(hipStream_t s, const ChainDesc* chains, int nchains, int nmax, VecRef dst) {
  hipLaunchKernelGGL(k_vec, dim3((nmax + 255) / 256, nchains), dim3(256), 0, s, chains, (int)VO_ONES, dst, dst);
  return hipGetLastError();
}

__global__ __launch_bounds__(256) void k_split_scales(const ChainDesc* chains, VecRef Dr, VecRef Br, VecRef Sr) {
  const ChainDesc& c = chains[blockIdx.y];
  const double* D = vecptr(c, Dr);
  double* ib = vecptr(c, Br);
  double* sm = vecptr(c, Sr);
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < c.n; i += gridDim.x * blockDim.x) {
    const double v = D[i];
    ib[i] = 1.0 / fmax(v, 1.0);
    sm[i] = fmin(v, 1.0);
  }
}
hipError_t launch_split_scales(hipStream_t s, const ChainDesc* chains, int nchains, int nmax, VecRef D,
                               VecRef inv_big, VecRef small) {
  hipLaunchKernelGGL(k_split_scales, dim3((nmax + 255) / 256, nchains), dim3(256), 0, s, chains, D, inv_big, small);
  return hipGetLastError();
}

// max|A - B| per chain: one block per chain, then fold into the drift slots.
// phase 0: clear the accumulator; phase 1: every block folds its max|A - B| into it with a
// 64-bit atomic max (non-negative doubles order like their bit patterns); phase 2: publish.
__global__ __launch_bounds__(256) void k_maxdiff(const ChainDesc* chains, MatRef Ar, MatRef Br, int phase) {
  const ChainDesc& c = chains[blockIdx.y];
  unsigned long long* acc = reinterpret_cast<unsigned long long*>(c.scal + S_DRIFT_TMP);
  if (!matptr(c, Ar)) return;  // a merged sweep's plan records no drift for this chain
  if (phase == 0) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *acc = 0ull;
    return;
  }
  if (phase == 2) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      const double t = __longlong_as_double((long long)*acc);
      c.scal[S_DRIFT_LAST] = t;
      c.scal[S_DRIFT_MAX] = fmax(c.scal[S_DRIFT_MAX], t);
    }
    return;
  }
  const size_t nn = (size_t)c.n * c.n;
  const cplx* A = matptr(c, Ar);
  const cplx* B = matptr(c, Br);
  double m = 0.0;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < nn; e += (size_t)gridDim.x * blockDim.x) {
    const cplx d = A[e] - B[e];
    m = fmax(m, sqrt(abs2(d)));
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off, 64));
  __shared__ double red[4];
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    const double t = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
    atomicMax(acc, (unsigned long long)__double_as_longlong(t));
  }
}
hipError_t launch_maxdiff(hipStream_t s, const ChainDesc* chains, int nchains, int nmax, MatRef A, MatRef B) {
  int blocks = (int)(((size_t)nmax * nmax + 255) / 256);
  if (blocks > 256) blocks = 256;
  hipLaunchKernelGGL(k_maxdiff, dim3(1, nchains), dim3(64), 0, s, chains, A, B, 0);
  hipLaunchKernelGGL(k_maxdiff, dim3(blocks, nchains), dim3(256), 0, s, chains, A, B, 1);
  hipLaunchKernelGGL(k_maxdiff, dim3(1, nchains), dim3(64), 0, s, chains, A, B, 2);
  return hipGetLastError();
}

// scal[slot] (+)= sum_i log(d_i)
__global__ __launch_bounds__(256) void k_sum_log(const ChainDesc* chains, VecRef dr, int slot, int accumulate) {
  const ChainDesc& c = chains[blockIdx.x];
  const double* d = vecptr(c, dr);
  double acc = 0.0;
  for (int i = threadIdx.x; i < c.n; i += blockDim.x) acc += log(d[i]);
  __shared__ double red[4];
  acc = block_sum(acc, red);
  if (threadIdx.x == 0) c.scal[slot] = (accumulate ? c.scal[slot] : 0.0) + acc;
}
hipError_t launch_sum_log(hipStream_t s, const ChainDesc* chains, int nchains, int, VecRef d, int slot,
                          int accumulate) {
  hipLaunchKernelGGL(k_sum_log, dim3(nchains), dim3(256), 0, s, chains, d, slot, accumulate);
  return hipGetLastError();
}

// scal[dst] += scal[src]
__global__ void k_add_scalar(const ChainDesc* chains, int dst, int src) {
  const ChainDesc& c = chains[blockIdx.x];
  if (threadIdx.x == 0) c.scal[dst] += c.scal[src];
}
hipError_t launch_add_scalar(hipStream_t s, const ChainDesc* chains, int nchains, int dst, int src) {
  hipLaunchKernelGGL(k_add_scalar, dim3(nchains), dim3(64), 0, s, chains, dst, src);
  return hipGetLastError();
}

__global__ void k_negate_scalar(const ChainDesc* chains, int slot) {
  const ChainDesc& c = chains[blockIdx.x];
  if (threadIdx.x == 0) c.scal[slot] = -c.scal[slot];
}
hipError_t launch_negate_scalar(hipStream_t s, const ChainDesc* chains, int nchains, int slot) {
  hipLaunchKernelGGL(k_negate_scalar, dim3(nchains), dim3(64), 0, s, chains, slot);
  return hipGetLastError();
}

}  // namespace dqmc
