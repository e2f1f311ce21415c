// Bosonic action S_B, its gradient, and the fermionic part of dS/dphi.
//
// Reference: calc_boson_action (action.jl:1-52). The gradient has no
// counterpart in the reference; it is the adjoint of
//   S(phi) = S_B(phi) - log det(1 + B_M ... B_1),
// whose fermionic part at slice l is, with B_l = H eV_l (eV rightmost,
// slice_matrices.jl:101-129) and G(l) = [1 + B_{l-1}..B_1 B_M..B_l]^-1 the
// stack's Green's function at slice l (stack.jl:336):
//   d log det / d phi_{a,i,l} = Re tr[(1 - G(l))_ii  eV_i^-1 d_a eV_i],
// a trace over the 4x4 flavour block of site i.
#include "chain.hpp"

namespace dqmc {

// one block per chain; S_B into scal[S_BOSON_ACTION]
__global__ __launch_bounds__(256) void k_boson_action(const ChainDesc* chains) {
  const ChainDesc& c = chains[blockIdx.x];
  const int L = c.L, N = c.N, M = c.M;
  double acc = 0.0;
  for (int e = threadIdx.x; e < N * M; e += blockDim.x) {
    const int l = e / N, i = e - l * N;
    const int y = i % L, x = i / L;
    const double* f = c.phi + 3 * (size_t)e;
    const double* ft = c.phi + 3 * ((size_t)i + (size_t)N * ((l + 1) % M));
    const double* fu = c.phi + 3 * ((size_t)((y + 1) % L + x * L) + (size_t)N * l);
    const double* fr = c.phi + 3 * ((size_t)(y + ((x + 1) % L) * L) + (size_t)N * l);
    const double sq = f[0] * f[0] + f[1] * f[1] + f[2] * f[2];
    if (!c.edrun) {
      double tt = 0.0, ts = 0.0;
      for (int a = 0; a < 3; ++a) {
        const double dt = f[a] - ft[a];
        const double du = f[a] - fu[a];
        const double dr = f[a] - fr[a];
        tt += dt * dt;
        ts += du * du + dr * dr;
      }
      acc += 0.5 / c.dtau / (c.c * c.c) * tt + c.dtau * 0.5 * ts;
      acc += c.dtau * c.r / 2.0 * sq + c.dtau * c.u / 4.0 * sq * sq;
    } else {
      acc += c.dtau * c.r / 2.0 * sq;
    }
  }
  __shared__ double red[4];
  acc = block_sum(acc, red);
  if (threadIdx.x == 0) c.scal[S_BOSON_ACTION] = acc;
}

hipError_t launch_boson_action(hipStream_t s, const ChainDesc* chains, int nchains, int) {
  hipLaunchKernelGGL(k_boson_action, dim3(nchains), dim3(256), 0, s, chains);
  return hipGetLastError();
}

// dS_B/dphi for every component: grad[a + 3(i + N l)] = (overwrites)
__global__ __launch_bounds__(256) void k_boson_grad(const ChainDesc* chains, double* const* grad) {
  const ChainDesc& c = chains[blockIdx.y];
  const int L = c.L, N = c.N, M = c.M;
  double* gr = grad[blockIdx.y];
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < N * M; e += gridDim.x * blockDim.x) {
    const int l = e / N, i = e - l * N;
    const int y = i % L, x = i / L;
    const double* f = c.phi + 3 * (size_t)e;
    const double* fp = c.phi + 3 * ((size_t)i + (size_t)N * ((l + 1) % M));
    const double* fm = c.phi + 3 * ((size_t)i + (size_t)N * ((l + M - 1) % M));
    const int nbs[4] = {(y + 1) % L + x * L, y + ((x + 1) % L) * L, (y + L - 1) % L + x * L,
                        y + ((x + L - 1) % L) * L};
    const double sq = f[0] * f[0] + f[1] * f[1] + f[2] * f[2];
    for (int a = 0; a < 3; ++a) {
      double gsum;
      if (!c.edrun) {
        double snb = 0.0;
        for (int k = 0; k < 4; ++k) snb += c.phi[3 * ((size_t)nbs[k] + (size_t)N * l) + a];
        gsum = (2.0 * f[a] - fp[a] - fm[a]) / (c.dtau * c.c * c.c) + c.dtau * (4.0 * f[a] - snb) +
               c.dtau * (c.r * f[a] + c.u * sq * f[a]);
      } else {
        gsum = c.dtau * c.r * f[a];
      }
      gr[3 * (size_t)e + a] = gsum;
    }
  }
}

hipError_t launch_boson_grad(hipStream_t s, const ChainDesc* chains, int nchains, int NMmax, double* const* grad) {
  int blocks = (NMmax + 255) / 256;
  if (blocks > 512) blocks = 512;
  hipLaunchKernelGGL(k_boson_grad, dim3(blocks, nchains), dim3(256), 0, s, chains, grad);
  return hipGetLastError();
}

// fermionic part at slice l (0-based) from the current G = G(l+1 in 1-based slices):
// grad[a,i,l] -= Re tr[(1 - G)_ii  eVop(phi,-1) d_a eVop(phi,+1)]
__global__ __launch_bounds__(256) void k_grad_slice(const ChainDesc* chains, int slice, double* const* grad) {
  const ChainDesc& c = chains[blockIdx.y];
  const int N = c.N, n = c.n;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const int l = slice;
  const cplx* G = c.mat[M_G];
  double* gr = grad[blockIdx.y];
  const double* f = c.phi + 3 * ((size_t)i + (size_t)N * l);
  const double ld = c.lam * c.dtau;
  const double nrm = sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
  const double x = ld * nrm;
  const double sh = sinh(x) / nrm, C = cosh(x);
  // (1 - G)_ii, row-major over flavour blocks
  cplx OG[16];
  for (int a = 0; a < 4; ++a)
    for (int b = 0; b < 4; ++b) {
      const cplx gv = G[(size_t)(i + b * N) * n + (i + a * N)];
      OG[a * 4 + b] = mk((a == b ? 1.0 : 0.0) - gv.re, -gv.im);
    }
  // eVop(phi, -1): C, S = -(i f1 - f0) sh, R = f2 sh
  auto block = [](double Cc, cplx S, double R, cplx* E) {
    const cplx z = mk(0, 0);
    E[0] = mk(Cc, 0);  E[1] = S;         E[2] = z;          E[3] = mk(R, 0);
    E[4] = conj(S);    E[5] = mk(Cc, 0); E[6] = mk(-R, 0);  E[7] = z;
    E[8] = z;          E[9] = mk(-R, 0); E[10] = mk(Cc, 0); E[11] = conj(S);
    E[12] = mk(R, 0);  E[13] = z;        E[14] = S;         E[15] = mk(Cc, 0);
  };
  cplx Einv[16];
  block(C, mk(f[0] * sh, -f[1] * sh), f[2] * sh, Einv);
  const double dsh_fac = (ld * C - sh) / (nrm * nrm);  // d sh / d phi_a = dsh_fac * phi_a
  const cplx St = mk(-f[0], f[1]);                      // S~ = i phi_1 - phi_0
  const double Rt = -f[2];
  for (int a = 0; a < 3; ++a) {
    const double dC = ld * sh * f[a];
    const double dsh = dsh_fac * f[a];
    cplx Sa = mk(dsh * St.re, dsh * St.im);
    double Ra = dsh * Rt;
    if (a == 0) Sa.re += -sh;
    if (a == 1) Sa.im += sh;
    if (a == 2) Ra += -sh;
    cplx dE[16], P[16];
    block(dC, Sa, Ra, dE);
    // P = Einv dE ; tr(OG P)
    for (int r = 0; r < 4; ++r)
      for (int q = 0; q < 4; ++q) {
        cplx acc = mk(0, 0);
        for (int k = 0; k < 4; ++k) cfma(acc, Einv[r * 4 + k], dE[k * 4 + q]);
        P[r * 4 + q] = acc;
      }
    double tr = 0.0;
    for (int r = 0; r < 4; ++r)
      for (int k = 0; k < 4; ++k) {
        const cplx t = OG[r * 4 + k] * P[k * 4 + r];
        tr += t.re;
      }
    gr[3 * ((size_t)i + (size_t)N * l) + a] -= tr;
  }
}

// y += a x over n doubles (HMC leapfrog kicks and drifts)
__global__ __launch_bounds__(256) void k_daxpy(double* y, const double* x, double a, long long n) {
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long long)gridDim.x * blockDim.x)
    y[e] = fma(a, x[e], y[e]);
}
hipError_t launch_daxpy(hipStream_t s, double* y, const double* x, double a, long long n) {
  long long blocks = (n + 255) / 256;
  if (blocks > 1024) blocks = 1024;
  hipLaunchKernelGGL(k_daxpy, dim3((unsigned)blocks), dim3(256), 0, s, y, x, a, n);
  return hipGetLastError();
}

hipError_t launch_grad_slice(hipStream_t s, const ChainDesc* chains, int nchains, int Nmax, int slice,
                             double* const* grad) {
  hipLaunchKernelGGL(k_grad_slice, dim3((Nmax + 255) / 256, nchains), dim3(256), 0, s, chains, slice, grad);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Global update (global_updates.jl:1-58): uniform shift of every component,
// decision, conditional restore. One chain per launch (chains[0]).
// ---------------------------------------------------------------------------

// phi_bak <- phi; phi[a, :, :] += -b + 2 b u_a with the next three uniforms
// (global_update_perform_shift!, global_updates.jl:10-16; randuniform, dqmc_framework.jl:628)
__global__ __launch_bounds__(256) void k_global_shift(const ChainDesc* chains, double* phi_bak, double box) {
  const ChainDesc& c = chains[0];
  __shared__ double sh[3];
  __shared__ int ok;
  if (threadIdx.x == 0) {
    const long long cur = *c.ucursor;
    ok = (cur + 3 <= c.n_uniforms) ? 1 : 0;
    if (ok) {
      for (int a = 0; a < 3; ++a) sh[a] = -box + 2 * box * c.uniforms[cur + a];
    } else {
      *c.error = 2;
    }
  }
  __syncthreads();
  if (!ok) return;
  const long long tot = 3ll * c.N * c.M;
  for (long long e = threadIdx.x; e < tot; e += blockDim.x) {
    const double v = c.phi[e];
    phi_bak[e] = v;
    c.phi[e] = v + sh[e % 3];
  }
  if (threadIdx.x == 0) *c.ucursor += 3;
}

// p_acc = exp(-(S_B new - S_B old)) exp(log_det old - log_det new); accept iff p_acc > 1 or
// u < p_acc with one more uniform (global_updates.jl:38-57). On rejection the scalars are restored.
__global__ void k_global_decide(const ChainDesc* chains) {
  const ChainDesc& c = chains[0];
  if (threadIdx.x != 0) return;
  double* sc = c.scal;
  const double S_old = sc[S_GB_ACTION], ld_old = sc[S_GB_LOGDET];
  const double p_boson = exp(-(sc[S_BOSON_ACTION] - S_old));
  const double p_fermion = exp(ld_old - sc[S_LOGDET]);
  const double p_acc = p_boson * p_fermion;
  int acc = 1;
  double u = -1.0;
  if (!(p_acc > 1.0)) {
    const long long cur = *c.ucursor;
    if (cur < c.n_uniforms) {
      u = c.uniforms[cur];
      *c.ucursor = cur + 1;
      acc = u < p_acc ? 1 : 0;
    } else {
      *c.error = 2;
      acc = 0;
    }
  }
  sc[S_GLOBAL_ACC] = acc;
  sc[S_GLOBAL_PACC] = p_acc;
  sc[S_GLOBAL_NACC] += acc;
  sc[S_GLOBAL_NPROP] += 1.0;
  if (u >= 0.0) sc[S_MINRATIO] = fmin(sc[S_MINRATIO], fabs(p_acc - u));
  if (!acc) {
    sc[S_BOSON_ACTION] = S_old;
    sc[S_LOGDET] = ld_old;
  }
}

// dst_k <- src_k for every listed buffer iff the last global update was rejected
__global__ __launch_bounds__(256) void k_global_restore(const double* flag, RestoreList r) {
  if (*flag != 0.0) return;
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (int k = 0; k < r.n; ++k) {
    double* dst = r.dst[k];
    const double* src = r.src[k];
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < r.count[k]; e += stride)
      dst[e] = src[e];
  }
}

hipError_t launch_global_shift(hipStream_t s, const ChainDesc* chain, double* phi_bak, double box) {
  hipLaunchKernelGGL(k_global_shift, dim3(1), dim3(256), 0, s, chain, phi_bak, box);
  return hipGetLastError();
}
hipError_t launch_global_decide(hipStream_t s, const ChainDesc* chain) {
  hipLaunchKernelGGL(k_global_decide, dim3(1), dim3(64), 0, s, chain);
  return hipGetLastError();
}
hipError_t launch_global_restore(hipStream_t s, const double* flag, const RestoreList& r) {
  hipLaunchKernelGGL(k_global_restore, dim3(1024), dim3(256), 0, s, flag, r);
  return hipGetLastError();
}

}  // namespace dqmc
