// Fermionic observables built on the equal-time and time-displaced Green's functions:
// equal-time and zero-frequency pairing correlations (measure_etpc!, measure_zfpc!,
// fermion_measurements.jl:388-470, 500-562), equal-time and zero-frequency charge-density
// correlations (measure_etcdc!, measure_zfcdc!, :592-657, :688-778) and zero-frequency
// current-current correlations (measure_zfccc!, :814-899). Every one is a sum over the
// imaginary times tau (or a single equal-time G), the origins r0 and the displacements r of
// a few Wick products of Green's-function entries, binned by r.
//
// Layout: the reference's L x L result arrays are indexed [y+1, x+1] (column-major), i.e.
// displacement r = (x, y) has linear index y + L x; a site (x, y) of the lattice is
// y + L x as well (sql = reshape(1:N, L, L), lattice.jl:109). One workgroup takes one tau
// and a chunk of origins; each thread owns displacements r = tid, tid + 256, ... and keeps
// their bins in registers, so nothing is shared; the partial bins of every (tau, chunk)
// are summed in a fixed order by a second kernel (deterministic results).
#include <algorithm>

#include "chain.hpp"

namespace dqmc {

namespace {

// 0-based flavours: xu, yd, xd, yu = 0, 1, 2, 3 (fermion_measurements.jl:14-18);
// flv(band, spin) = band + spin with X, Y = 1, 2 and UP, DOWN = 0, 2 (1-based), so
// (X, UP) = 0, (Y, UP) = 1, (X, DOWN) = 2, (Y, DOWN) = 3.
__device__ inline int flv(int band, int spin) { return band + spin; }  // band 0 = X, 1 = Y; spin 0 / 2

__device__ inline int site(int L, int x, int y) {
  x %= L;
  y %= L;
  return y + L * x;
}

__device__ inline cplx gat(const cplx* G, int n, int i, int j) { return ld(gp(G) + (size_t)j * n + i); }

// pairing: Re[G(i1,i4) G(i2,i3) - G(i1,i3) G(i2,i4)] (_pc_ev, fermion_measurements.jl:442-449)
__device__ inline double pc_ev(const cplx* G, int n, int N, int r1, int r2, int r3, int r4, int j1, int j2, int j3,
                               int j4) {
  const int i1 = j1 * N + r1, i2 = j2 * N + r2, i3 = j3 * N + r3, i4 = j4 * N + r4;
  const cplx a = gat(G, n, i1, i4) * gat(G, n, i2, i3);
  const cplx b = gat(G, n, i1, i3) * gat(G, n, i2, i4);
  return a.re - b.re;
}

// charge density (_etcdc_ev :641-653 with ET = true, _zfcdc_ev :753-766 with ET = false)
template <bool ET>
__device__ inline cplx cdc_ev(const cplx* Gtau, const cplx* G0, const cplx* Gt0, const cplx* G0t, int n, int N,
                              int r1, int r2, int r3, int r4, int j1, int j2, int j3, int j4) {
  const int i1 = j1 * N + r1, i2 = j2 * N + r2, i3 = j3 * N + r3, i4 = j4 * N + r4;
  cplx a = -gat(Gtau, n, i2, i1);
  if (i1 == i2) a.re += 1.0;
  cplx b = -gat(G0, n, i4, i3);
  if (i3 == i4) b.re += 1.0;
  cplx c1, c2;
  if (ET) {
    c1 = -gat(Gtau, n, i4, i1);
    if (i1 == i4) c1.re += 1.0;
    c2 = gat(Gtau, n, i2, i3);
  } else {
    c1 = -gat(G0t, n, i4, i1);
    c2 = gat(Gt0, n, i2, i3);
  }
  return a * b + c1 * c2;
}

}  // namespace

// KIND 0: pairing (2 real bins: minus, plus), 1: charge density (2 complex bins: minus, plus),
// 2: current-current (2 complex bins: xx, yy). ET: single equal-time G (Gtau = G0 = Gt0,
// stride 0, one tau). part: [gridDim.x * gridDim.y][2][N] complex partial bins.
template <int KIND, bool ET>
__global__ __launch_bounds__(256) void k_fermion_obs(int L, const cplx* Gtau, const cplx* G0, const cplx* Gt0,
                                                     size_t sGt0, const cplx* G0t, size_t sG0t, const cplx* tnb,
                                                     cplx* part) {
  const int N = L * L, n = 4 * N;
  const int tau = blockIdx.x, chunk = blockIdx.y, nch = gridDim.y;
  const cplx* Gt = Gt0 + sGt0 * tau;
  const cplx* Gz = G0t + sG0t * tau;
  const int o0 = (int)(((long long)N * chunk) / nch), o1 = (int)(((long long)N * (chunk + 1)) / nch);
  cplx* out = part + ((size_t)blockIdx.y * gridDim.x + blockIdx.x) * 2 * N;
  for (int r = threadIdx.x; r < N; r += blockDim.x) {
    const int x = r / L, y = r - x * L;
    cplx b0 = mk(0, 0), b1 = mk(0, 0);
    for (int o = o0; o < o1; ++o) {
      const int X0 = o / L, Y0 = o - X0 * L;
      const int ri = site(L, X0 + x, Y0 + y), r0 = site(L, X0, Y0);
      if (KIND == 0) {
        // js = ((xd,xu,xu,xd), (xd,xu,yu,yd), (yd,yu,xu,xd), (yd,yu,yu,yd)) (:515-516)
        const double e1 = pc_ev(Gt, n, N, ri, ri, r0, r0, 2, 0, 0, 2);
        const double e2 = pc_ev(Gt, n, N, ri, ri, r0, r0, 2, 0, 3, 1);
        const double e3 = pc_ev(Gt, n, N, ri, ri, r0, r0, 1, 3, 0, 2);
        const double e4 = pc_ev(Gt, n, N, ri, ri, r0, r0, 1, 3, 3, 1);
        b0.re += e1 - e2 - e3 + e4;
        b1.re += e1 + e2 + e3 + e4;
      } else if (KIND == 1) {
        for (int s1 = 0; s1 <= 2; s1 += 2)
          for (int s2 = 0; s2 <= 2; s2 += 2) {
            const int xs1 = flv(0, s1), ys1 = flv(1, s1), xs2 = flv(0, s2), ys2 = flv(1, s2);
            const cplx e1 = cdc_ev<ET>(Gtau, G0, Gt, Gz, n, N, ri, ri, r0, r0, xs1, xs1, xs2, xs2);
            const cplx e2 = cdc_ev<ET>(Gtau, G0, Gt, Gz, n, N, ri, ri, r0, r0, xs1, xs1, ys2, ys2);
            const cplx e3 = cdc_ev<ET>(Gtau, G0, Gt, Gz, n, N, ri, ri, r0, r0, ys1, ys1, xs2, xs2);
            const cplx e4 = cdc_ev<ET>(Gtau, G0, Gt, Gz, n, N, ri, ri, r0, r0, ys1, ys1, ys2, ys2);
            b0 += (e1 - e2) - e3 + e4;
            b1 += e1 + e2 + e3 + e4;
          }
      } else {
        // tnb[d][f][i][0/1] = T[fN + i, fN + j], T[fN + j, fN + i] with j = i + x^ (d = 0) or
        // i + y^ (d = 1): the hopping entries the current operator needs (:845-851)
        for (int d = 0; d < 2; ++d) {
          const int rj = d == 0 ? site(L, X0 + x + 1, Y0 + y) : site(L, X0 + x, Y0 + y + 1);
          const int r0p = d == 0 ? site(L, X0 + 1, Y0) : site(L, X0, Y0 + 1);
          cplx lam = mk(0, 0);
          for (int a1 = 0; a1 < 2; ++a1)
            for (int a2 = 0; a2 < 2; ++a2)
              for (int s1 = 0; s1 <= 2; s1 += 2)
                for (int s2 = 0; s2 <= 2; s2 += 2) {
                  const int f1 = flv(a1, s1), f2 = flv(a2, s2);
                  const int ais = f1 * N + ri, ajs = f1 * N + rj, ap0sp = f2 * N + r0, ap0psp = f2 * N + r0p;
                  const cplx* t1 = tnb + 2 * ((size_t)(d * 4 + f1) * N + ri);
                  const cplx* t2 = tnb + 2 * ((size_t)(d * 4 + f2) * N + r0);
                  const cplx tij = ld(gp(t1)), tji = ld(gp(t1) + 1), t00p = ld(gp(t2)), t0p0 = ld(gp(t2) + 1);
                  // uncorrelated part
                  const cplx u1 = tij * gat(Gtau, n, ajs, ais) - tji * gat(Gtau, n, ais, ajs);
                  const cplx u2 = t00p * gat(G0, n, ap0psp, ap0sp) - t0p0 * gat(G0, n, ap0sp, ap0psp);
                  lam += u1 * u2;
                  // correlated part
                  lam += -(tij * t00p) * gat(Gz, n, ap0psp, ais) * gat(Gt, n, ajs, ap0sp);
                  lam += (tij * t0p0) * gat(Gz, n, ap0sp, ais) * gat(Gt, n, ajs, ap0psp);
                  lam += (tji * t00p) * gat(Gz, n, ap0psp, ajs) * gat(Gt, n, ais, ap0sp);
                  lam += -(tji * t0p0) * gat(Gz, n, ap0sp, ajs) * gat(Gt, n, ais, ap0psp);
                }
          if (d == 0)
            b0 += lam;
          else
            b1 += lam;
        }
      }
    }
    st(gp(out) + r, b0);
    st(gp(out) + N + r, b1);
  }
}

// res[k][r] = scale * sum over the partial blocks (fixed order) of part[b][k][r]
__global__ __launch_bounds__(256) void k_fermion_obs_reduce(int N, int nblocks, const cplx* part, double scale,
                                                            cplx* res) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= 2 * N) return;
  cplx acc = mk(0, 0);
  for (int b = 0; b < nblocks; ++b) acc += ld(gp(part) + (size_t)b * 2 * N + e);
  st(gp(res) + e, scale * acc);
}

int fermion_obs_blocks(int ntau, int N) {
  const int chunks = std::max(1, std::min(N, (512 + ntau - 1) / ntau));
  return ntau * chunks;
}

hipError_t launch_fermion_obs(hipStream_t s, int kind, int L, int ntau, const cplx* Gtau, const cplx* G0,
                              const cplx* Gt0, size_t sGt0, const cplx* G0t, size_t sG0t, const cplx* tnb, cplx* part,
                              cplx* res) {
  const int N = L * L;
  if (kind < 0 || kind > 2 || ntau < 1 || L < 1) return hipErrorInvalidValue;
  const int chunks = fermion_obs_blocks(ntau, N) / ntau;
  const bool et = sGt0 == 0;
  dim3 grid(ntau, chunks);
  if (kind == 0)
    hipLaunchKernelGGL((k_fermion_obs<0, false>), grid, dim3(256), 0, s, L, Gtau, G0, Gt0, sGt0, G0t, sG0t, tnb,
                       part);
  else if (kind == 1 && et)
    hipLaunchKernelGGL((k_fermion_obs<1, true>), grid, dim3(256), 0, s, L, Gtau, G0, Gt0, sGt0, G0t, sG0t, tnb,
                       part);
  else if (kind == 1)
    hipLaunchKernelGGL((k_fermion_obs<1, false>), grid, dim3(256), 0, s, L, Gtau, G0, Gt0, sGt0, G0t, sG0t, tnb,
                       part);
  else
    hipLaunchKernelGGL((k_fermion_obs<2, false>), grid, dim3(256), 0, s, L, Gtau, G0, Gt0, sGt0, G0t, sG0t, tnb,
                       part);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  // r0 mean (/ N); the current-current correlations carry the overall minus sign of i^2 (:895)
  const double scale = kind == 2 ? -1.0 / N : 1.0 / N;
  hipLaunchKernelGGL(k_fermion_obs_reduce, dim3((2 * N + 255) / 256), dim3(256), 0, s, N, ntau * chunks,
                     (const cplx*)part, scale, res);
  return hipGetLastError();
}

}  // namespace dqmc
