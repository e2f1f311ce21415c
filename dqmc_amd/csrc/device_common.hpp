// Device-side building blocks shared by every DQMC kernel: complex FP64
// arithmetic, wave/block reductions, and the inter-workgroup barrier used by a
// "chain group" (the set of workgroups that cooperate on one Markov chain).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <map>
#include <mutex>
#include <type_traits>
#include <utility>

namespace dqmc {

// Raises a kernel's dynamic-LDS limit to at least `lds` bytes on the current device. The
// attribute is per device and per function; the cache of what has been set is shared by
// every context and thread of the process, so it is guarded by a mutex.
inline hipError_t ensure_dynamic_lds(const void* fn, size_t lds) {
  if (lds <= 65536) return hipSuccess;
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  static std::mutex mu;
  static std::map<std::pair<int, const void*>, size_t> set;
  std::lock_guard<std::mutex> lock(mu);
  size_t& cur = set[std::make_pair(dev, fn)];
  if (lds <= cur) return hipSuccess;
  e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e == hipSuccess) cur = lds;
  return e;
}

struct cplx {
  double re, im;
};

__host__ __device__ inline cplx mk(double r, double i) { return cplx{r, i}; }
__host__ __device__ inline cplx operator+(cplx a, cplx b) { return cplx{a.re + b.re, a.im + b.im}; }
__host__ __device__ inline cplx operator-(cplx a, cplx b) { return cplx{a.re - b.re, a.im - b.im}; }
__host__ __device__ inline cplx operator-(cplx a) { return cplx{-a.re, -a.im}; }
__host__ __device__ inline cplx operator*(cplx a, cplx b) {
  return cplx{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
__host__ __device__ inline cplx operator*(double s, cplx a) { return cplx{s * a.re, s * a.im}; }
__host__ __device__ inline cplx conj(cplx a) { return cplx{a.re, -a.im}; }
__host__ __device__ inline double abs2(cplx a) { return a.re * a.re + a.im * a.im; }
__host__ __device__ inline cplx& operator+=(cplx& a, cplx b) {
  a.re += b.re;
  a.im += b.im;
  return a;
}
__host__ __device__ inline cplx& operator-=(cplx& a, cplx b) {
  a.re -= b.re;
  a.im -= b.im;
  return a;
}
// a += b*c
__host__ __device__ inline void cfma(cplx& a, cplx b, cplx c) {
  a.re = fma(b.re, c.re, a.re);
  a.re = fma(-b.im, c.im, a.re);
  a.im = fma(b.re, c.im, a.im);
  a.im = fma(b.im, c.re, a.im);
}
// a += conj(b)*c
__host__ __device__ inline void cfma_cj(cplx& a, cplx b, cplx c) {
  a.re = fma(b.re, c.re, a.re);
  a.re = fma(b.im, c.im, a.re);
  a.im = fma(b.re, c.im, a.im);
  a.im = fma(-b.im, c.re, a.im);
}
// 1/a (Smith's algorithm, avoids overflow in |a|^2)
__host__ __device__ inline cplx cinv(cplx a) {
  if (fabs(a.re) >= fabs(a.im)) {
    double r = a.im / a.re, d = a.re + a.im * r;
    return cplx{1.0 / d, -r / d};
  }
  double r = a.re / a.im, d = a.re * r + a.im;
  return cplx{r / d, -1.0 / d};
}

// Global-memory views. Buffer pointers read from a descriptor are generic, so accesses
// through them compile to FLAT instructions, which count in LGKM_CNT as well as VM_CNT:
// every wait for an LDS result, and every barrier, then also waits for the outstanding
// global stores. Through these address_space(1) views the same accesses compile to GLOBAL
// instructions, which only count in VM_CNT.
typedef __attribute__((address_space(1))) cplx gcplx;
typedef __attribute__((address_space(1))) double gdouble;
typedef __attribute__((address_space(1))) int gint;
__device__ inline gcplx* gp(cplx* p) { return (gcplx*)p; }
__device__ inline const gcplx* gp(const cplx* p) { return (const gcplx*)p; }
__device__ inline gdouble* gp(double* p) { return (gdouble*)p; }
__device__ inline const gdouble* gp(const double* p) { return (const gdouble*)p; }
__device__ inline gint* gp(int* p) { return (gint*)p; }
__device__ inline const gint* gp(const int* p) { return (const gint*)p; }
__device__ inline cplx ld(const gcplx* p) { return mk(p->re, p->im); }
__device__ inline void st(gcplx* p, cplx v) {
  p->re = v.re;
  p->im = v.im;
}
// streaming store: for data the next reader takes from another XCD (less for the next
// agent-scope release to write back from this XCD's L2)
__device__ inline void st_nt(gcplx* p, cplx v) {
  __builtin_nontemporal_store(v.re, &p->re);
  __builtin_nontemporal_store(v.im, &p->im);
}

// ---------------------------------------------------------------------------
// Reductions (wave64)
//
// wave_sum: inclusive scans inside each row of 16 lanes with DPP row shifts (zero fill),
// then the four row totals (lanes 15, 31, 47, 63) are read into scalar registers and
// added. No LDS traffic; every lane returns the same value. Needs all 64 lanes active.
// ---------------------------------------------------------------------------
template <int CTRL>
__device__ inline double dpp_move(double x) {
  const long long b = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffffll), CTRL, 0xF, 0xF, true);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xF, 0xF, true);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned int)lo);
}
__device__ inline double read_lane(double x, int lane) {
  const long long b = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), lane);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned int)lo);
}
__device__ inline double wave_sum(double v) {
  v += dpp_move<0x111>(v);  // row_shr:1
  v += dpp_move<0x112>(v);  // row_shr:2
  v += dpp_move<0x114>(v);  // row_shr:4
  v += dpp_move<0x118>(v);  // row_shr:8
  return (read_lane(v, 15) + read_lane(v, 31)) + (read_lane(v, 47) + read_lane(v, 63));
}
__device__ inline cplx wave_sum(cplx v) {
  v.re = wave_sum(v.re);
  v.im = wave_sum(v.im);
  return v;
}

// static_for<I, N>(f): f(std::integral_constant<int, k>{}) for k = I .. N-1, each k a
// compile-time constant in f (register arrays indexed by it stay in registers).
template <int I, int N, class F>
__device__ inline void static_for(F&& f) {
  if constexpr (I < N) {
    f(std::integral_constant<int, I>{});
    static_for<I + 1, N>(f);
  }
}

// Sums of many values at once. v_permlane32_swap exchanges the upper half of x with the
// lower half of y (v_permlane16_swap: the odd rows of x with the even rows of y), so
// swap_sum<32>(a, b) = a[l] + a[l ^ 32] on lanes 0-31 and b[l] + b[l ^ 32] on lanes 32-63
// (swap_sum<16>: a on rows 0 and 2, b on rows 1 and 3, each summed with lane l ^ 16).
template <int S>
__device__ inline double swap_sum(double a, double b) {
  const long long ab = __double_as_longlong(a), bb = __double_as_longlong(b);
  const unsigned alo = (unsigned)(ab & 0xffffffffll), ahi = (unsigned)(ab >> 32);
  const unsigned blo = (unsigned)(bb & 0xffffffffll), bhi = (unsigned)(bb >> 32);
  const auto lo = S == 32 ? __builtin_amdgcn_permlane32_swap(alo, blo, false, false)
                          : __builtin_amdgcn_permlane16_swap(alo, blo, false, false);
  const auto hi = S == 32 ? __builtin_amdgcn_permlane32_swap(ahi, bhi, false, false)
                          : __builtin_amdgcn_permlane16_swap(ahi, bhi, false, false);
  const double r0 = __longlong_as_double(((long long)hi[0] << 32) | (long long)lo[0]);
  const double r1 = __longlong_as_double(((long long)hi[1] << 32) | (long long)lo[1]);
  return r0 + r1;
}
// One butterfly stage inside 16-lane rows: the lane pairs l, l ^ M (CTRL is the DPP
// control that reads lane l ^ M: row_mirror 15, row_half_mirror 7, quad_perm 3 and 1).
// With c >= 1 values left per lane, the lane whose bit B is clear keeps the first half
// and receives its partner's copy of it, the other keeps the second half; with one
// value left the stage is a plain all-reduce of the pair.
template <int CTRL, int B, int C>
__device__ inline void row_stage(double (&v)[C], int lane) {
  if constexpr (C == 1) {
    v[0] += dpp_move<CTRL>(v[0]);
  } else {
    const bool hi = (lane >> B) & 1;
#pragma unroll
    for (int k = 0; k < C / 2; ++k) {
      const double a = v[k], b = v[k + C / 2];
      v[k] = (hi ? b : a) + dpp_move<CTRL>(hi ? a : b);
    }
  }
}
// wave_reduce_scatter: every one of v[0..V-1] (V = 8, 16, 32 or 64) summed over the 64 lanes
// of the wave by a butterfly reduce-scatter, each stage halving the values a lane carries,
// so the whole reduction costs about 2 V exchanges instead of 12 V for V separate
// wave_sums. On return lane l holds the total of value l >> (6 - log2 V) in v[0] (lanes
// that share it hold the same bits). Needs all 64 lanes active.
template <int V>
__device__ inline void wave_reduce_scatter(double (&v)[V]) {
  static_assert(V == 8 || V == 16 || V == 32 || V == 64, "V");
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int k = 0; k < V / 2; ++k) v[k] = swap_sum<32>(v[k], v[k + V / 2]);
#pragma unroll
  for (int k = 0; k < V / 4; ++k) v[k] = swap_sum<16>(v[k], v[k + V / 4]);
  double (&w8)[V / 4] = reinterpret_cast<double (&)[V / 4]>(v);
  row_stage<0x140, 3>(w8, lane);  // row_mirror: l ^ 15
  constexpr int C4 = V / 8 > 1 ? V / 8 : 1;
  double (&w4)[C4] = reinterpret_cast<double (&)[C4]>(v);
  row_stage<0x141, 2>(w4, lane);  // row_half_mirror: l ^ 7
  constexpr int C2 = V / 16 > 1 ? V / 16 : 1;
  double (&w2)[C2] = reinterpret_cast<double (&)[C2]>(v);
  row_stage<0x1B, 1>(w2, lane);  // quad_perm [3,2,1,0]: l ^ 3
  constexpr int C1 = V / 32 > 1 ? V / 32 : 1;
  double (&w1)[C1] = reinterpret_cast<double (&)[C1]>(v);
  row_stage<0xB1, 0>(w1, lane);  // quad_perm [1,0,3,2]: l ^ 1
}

// wave_sum of each of NB complex values (NB = 4, 8, 16 or 32): one reduce-scatter, then the
// totals are read back into every lane (scalar registers)
template <int NB>
__device__ inline void wave_sum_many(cplx (&w)[NB]) {
  constexpr int V = 2 * NB;
  constexpr int LS = V == 8 ? 3 : V == 16 ? 2 : V == 32 ? 1 : 0;  // lane l holds value l >> LS
  double v[V];
#pragma unroll
  for (int j = 0; j < NB; ++j) {
    v[2 * j] = w[j].re;
    v[2 * j + 1] = w[j].im;
  }
  wave_reduce_scatter<V>(v);
#pragma unroll
  for (int j = 0; j < NB; ++j) w[j] = mk(read_lane(v[0], (2 * j) << LS), read_lane(v[0], (2 * j + 1) << LS));
}

// XCD-aware workgroup numbering. The dispatcher hands consecutive workgroups of a launch
// to the 8 XCDs in turn, and each XCD has its own L2. Renumbering so that every XCD gets a
// contiguous range of logical blocks puts the workgroups of one chain (blockIdx.y) on as
// few XCDs as possible, where they share the chain's data in L2. A pure relabelling of a
// (gridDim.x, gridDim.y) grid: every logical block is run exactly once.
constexpr int kNumXcd = 8;
__device__ inline void xcd_block(int& bx, int& by) {
  const int X = gridDim.x, total = X * gridDim.y;
  const int pid = blockIdx.y * X + blockIdx.x;
  const int per = total / kNumXcd, rem = total % kNumXcd;
  const int k = pid % kNumXcd, slot = pid / kNumXcd;
  const int l = k * per + min(k, rem) + slot;
  by = l / X;
  bx = l - by * X;
}

// Block-wide sum; `red` is LDS scratch of >= blockDim.x/64 doubles (x2 for cplx).
__device__ inline double block_sum(double v, double* red) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
  v = wave_sum(v);
  __syncthreads();
  if (lane == 0) red[wid] = v;
  __syncthreads();
  double t = 0.0;
  for (int i = 0; i < nw; ++i) t += red[i];
  return t;
}

// ---------------------------------------------------------------------------
// Chain-group barrier
//
// A chain group is a set of `nblocks` workgroups cooperating on one chain,
// guaranteed co-resident by a cooperative launch. The barrier is a generation
// counter in device memory (agent-scope atomics, release/acquire fences).
// Every wait is bounded: when the spin limit is hit the group's error word is
// set and every later barrier returns immediately, so a logic error ends the
// kernel instead of hanging the device.
// ---------------------------------------------------------------------------
struct GroupSync {
  unsigned int* count;  // arrivals in the current generation
  unsigned int* gen;    // generation counter
  int* error;           // nonzero once a wait timed out
  unsigned int nblocks;
};

constexpr long long kSpinLimit = 1ll << 26;  // ~ seconds of polling before giving up

__device__ inline bool group_failed(const GroupSync& g) {
  return __hip_atomic_load(g.error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0;
}

__device__ inline void group_sync(const GroupSync& g) {
  if (g.nblocks <= 1) {
    __syncthreads();
    return;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    // Read the generation before arriving; the arrival is a release (publishes this
    // workgroup's writes). Polling uses relaxed loads so the wait does not invalidate
    // caches on every iteration; one acquire fence after the wait makes the other
    // workgroups' writes visible.
    const unsigned int my_gen = __hip_atomic_load(g.gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const unsigned int arrived =
        __hip_atomic_fetch_add(g.count, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT) + 1u;
    if (arrived == g.nblocks) {
      __hip_atomic_store(g.count, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(g.gen, my_gen + 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    } else {
      // The error word is read every 32nd poll only: each agent-scope load goes past the
      // XCD's L2, and a second load per poll would double the time between polls of gen.
      long long spins = 0;
      while (__hip_atomic_load(g.gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == my_gen) {
        if ((++spins & 31) == 0) {
          if (__hip_atomic_load(g.error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) break;
          if (spins > kSpinLimit) {
            __hip_atomic_store(g.error, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            break;
          }
        }
        __builtin_amdgcn_s_sleep(1);
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  }
  __syncthreads();
}

}  // namespace dqmc
