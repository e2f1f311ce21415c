// C ABI and host orchestration of the DQMC sweep path.
//
// The host side mirrors the reference's control flow (stack.jl: build_stack,
// add_slice_sequence_left/right, calculate_greens, calculate_logdet,
// wrap_greens!, propagate; local_updates.jl: local_updates) and enqueues the
// device kernels on one HIP stream per shape group; no call synchronises unless
// it returns data to the host. A single-chain context is a batch of one chain.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <random>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/dqmc_hip.h"
#include "chain.hpp"

using namespace dqmc;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define HIPCHK(expr)                                                                                  \
  do {                                                                                                \
    hipError_t _e = (expr);                                                                           \
    if (_e != hipSuccess)                                                                             \
      return fail(DQMC_E_HIP, std::string(#expr) + ": " + hipGetErrorString(_e));                     \
  } while (0)

#define RC(x)            \
  do {                   \
    int _rc = (x);       \
    if (_rc) return _rc; \
  } while (0)

constexpr int kMaxWG = 128;

MatRef NM(int id) { return MatRef{MK_NAMED, id}; }
MatRef US(int k) { return MatRef{MK_USTACK, k}; }
MatRef TS(int k) { return MatRef{MK_TSTACK, k}; }
VecRef NV(int id) { return VecRef{VK_NAMED, id}; }
VecRef DS(int k) { return VecRef{VK_DSTACK, k}; }
const VecRef NOVEC{-1, 0};

struct ChainMem {
  dqmc_params p;
  int L, N, n, M, nelem;
  ChainDesc d;  // host copy of the device descriptor
  std::vector<void*> allocs;
  double* uniforms = nullptr;
  long long ucap = 0;
  double* log_buf = nullptr;
  long long log_cap = 0;
  // gradient / HMC buffers (allocated on first use)
  double* grad = nullptr;
  double** grad_ptr = nullptr;
  double* mom = nullptr;
  double* phi_bak = nullptr;
  // time-displaced Green's function storage (dqmc_tdgf_alloc)
  cplx* tdgf_mat = nullptr;
  double* tdgf_vec = nullptr;
  // global-update backup (gb_u/d/t_stack, gb_greens, gb_hsfield; stack.jl:134-143), one block
  double* gb = nullptr;
  // fermionic measurements (allocated on first use): the equal-time Green's function the
  // observables read (etgf), the hopping entries of the current operator, partial bins, results
  cplx* meas = nullptr;
  cplx* meas_etgf = nullptr;
  cplx* meas_tnb = nullptr;
  cplx* meas_part = nullptr;
  cplx* meas_res = nullptr;
  bool has_etgf = false, has_hopping = false;
};

struct Group {
  int n, N, L, M, s, nelem, chkr;
  int ngroups = 0;  // bond groups of the generic checkerboard (0: Assaad or dense)
  bool all_checks;
  std::vector<int> chains;  // indices into Batch::chains
  ChainDesc* d_desc = nullptr;
  hipStream_t stream = nullptr;
  int cur = 0, dir = 1;  // reference current_slice (1-based), direction
  int wg_qr = 1, wg_inv = 1, wg_local = 1;
  int greens_route = 0;  // 0: D-scale separation, 1: verbatim reference route
  int local_variant = 0;  // 0 auto, 1 G in global memory, 2 LDS-resident G tiles, 3 delayed updates,
                          // 4 windowed delayed updates
  int udt_route = 0;      // 0: blocked QR with pre-pivoting, 1: column-pivoted QR (zgeqp3 semantics)
  int qr_nb = 0;          // panel width of the blocked QR (0: the default, 8)
  // merged sweeps: per-block map of the blocked QR's 1-D grid (workgroups per chain in
  // proportion to n^3; launch_qr_blocked), nullptr for the uniform (wg_qr, chains) grid
  const int* qr_map = nullptr;
  int qr_map_blocks = 0;
  // merged sweeps: the same for the window-variant local updates (local_block_map)
  const int* local_map = nullptr;
  int local_map_blocks = 0;
  bool recomputed = false;
  // one sweep (M update steps) captured as a HIP graph per starting state and settings
  struct SweepGraph {
    int cur0, dir0, dolog;
    double box;
    int cur1, dir1;
    hipGraphExec_t exec;
  };
  std::vector<SweepGraph> graphs;
  void drop_graphs() {
    for (auto& e : graphs) (void)hipGraphExecDestroy(e.exec);
    graphs.clear();
  }
  // kernel-class timing (dqmc_profile)
  unsigned prof_mask = 0;
  struct Rec {
    int cls;
    hipEvent_t a, b;
  };
  std::vector<Rec> prof;
  std::vector<hipEvent_t> event_pool;
  hipEvent_t take_event() {
    if (!event_pool.empty()) {
      hipEvent_t e = event_pool.back();
      event_pool.pop_back();
      return e;
    }
    hipEvent_t e = nullptr;
    (void)hipEventCreate(&e);
    return e;
  }
  int nch() const { return (int)chains.size(); }
};

}  // namespace

struct dqmc_batch {
  int device = 0;
  std::vector<ChainMem> chains;
  std::vector<Group> groups;
  std::vector<int> chain_group;  // chain -> group index
  double* d_scratch = nullptr;   // small host-visible staging (detratio etc.)
  bool use_graph = false;
  // merged sweeps of several shape groups (chains of any L and M, one launch per kernel)
  bool merged = true;
  hipStream_t mstream = nullptr;
  int* d_mplan = nullptr;        // per-chain plan rows, then per-launch selector / slice rows
  size_t mplan_cap = 0;          // ints
  ChainDesc* d_mdesc = nullptr;  // descriptors of the active sets of a merged call
  size_t mdesc_cap = 0;          // descriptors
  std::vector<int> h_rows;  // host side of the plan (kept until copied)
  std::vector<dqmc::ChainDesc> h_mdesc;
  // merged sweeps replayed from HIP graphs, one per start state (dqmc_batch_use_graph)
  struct MergedGraph {
    std::vector<int> key, end;  // groups' (current_slice, direction) before / after
    std::vector<dqmc::ChainDesc> descs;  // the chains' descriptors the graph was captured with
    double box;
    int dolog;
    int* d_plan = nullptr;
    dqmc::ChainDesc* d_desc = nullptr;
    hipGraphExec_t exec = nullptr;
  };
  std::vector<MergedGraph> mgraphs;
  void drop_mgraphs() {
    if (!mgraphs.empty() && mstream) (void)hipStreamSynchronize(mstream);
    for (auto& e : mgraphs) {
      (void)hipGraphExecDestroy(e.exec);
      (void)hipFree(e.d_plan);
      (void)hipFree(e.d_desc);
    }
    mgraphs.clear();
  }
};

struct dqmc_ctx {
  dqmc_batch* b;
};

namespace {

int alloc(ChainMem& cm, void** p, size_t bytes) {
  hipError_t e = hipMalloc(p, bytes == 0 ? 16 : bytes);
  if (e != hipSuccess) return fail(DQMC_E_HIP, std::string("hipMalloc: ") + hipGetErrorString(e));
  cm.allocs.push_back(*p);
  return 0;
}

int validate(const dqmc_params* p, const dqmc_operators* o) {
  if (!p || !o) return fail(DQMC_E_ARG, "null params/operators");
  if (p->L < 2 || p->slices < 1 || p->safe_mult < 1) return fail(DQMC_E_ARG, "bad L/slices/safe_mult");
  if (p->slices % p->safe_mult != 0) return fail(DQMC_E_ARG, "slices must be a multiple of safe_mult");
  if (4 * p->L * p->L > 1024) return fail(DQMC_E_ARG, "n = 4L^2 must be <= 1024");
  if (p->checkerboard < 0 || p->checkerboard > 2) return fail(DQMC_E_ARG, "checkerboard must be 0, 1 or 2");
  if (p->checkerboard == 1 && (p->L % 2 != 0)) return fail(DQMC_E_ARG, "the Assaad checkerboard needs even L");
  if (!o->neighbors) return fail(DQMC_E_ARG, "neighbors required");
  if (p->checkerboard == 1 && (!o->plaq_sites || !o->plaq || !o->expmu))
    return fail(DQMC_E_ARG, "checkerboard operators required");
  if (p->checkerboard == 2) {
    if (!o->group_offsets || !o->bond_sites || !o->bond_blocks || !o->expmu)
      return fail(DQMC_E_ARG, "generic checkerboard operators required");
    if (o->n_groups < 1 || o->n_groups > kMaxBondGroups)
      return fail(DQMC_E_ARG, "generic checkerboard: 1 to 16 bond groups");
    const int N = p->L * p->L;
    if (o->group_offsets[0] != 0) return fail(DQMC_E_ARG, "generic checkerboard: group_offsets[0] must be 0");
    for (int g = 0; g < o->n_groups; ++g)
      if (o->group_offsets[g + 1] < o->group_offsets[g]) return fail(DQMC_E_ARG, "group_offsets must not decrease");
    const int nb = o->group_offsets[o->n_groups];
    for (int g = 0; g < o->n_groups; ++g) {
      std::vector<char> used(N, 0);
      for (int q = o->group_offsets[g]; q < o->group_offsets[g + 1]; ++q)
        for (int k = 0; k < 2; ++k) {
          const int st = o->bond_sites[2 * q + k];
          if (st < 0 || st >= N) return fail(DQMC_E_ARG, "generic checkerboard: bond site out of range");
          if (used[st]++) return fail(DQMC_E_ARG, "generic checkerboard: a group touches a site twice");
        }
    }
    if (nb < 1) return fail(DQMC_E_ARG, "generic checkerboard: no bonds");
  }
  if (!p->checkerboard && (!o->hop_exp || !o->hop_exp_inv)) return fail(DQMC_E_ARG, "dense operators required");
  return 0;
}

// complex n x n product on the host for setup only (exp(-dtau T/2)^2)
std::vector<double> host_cmatmul(const double* A, const double* B, int n, bool adjA) {
  std::vector<double> C((size_t)2 * n * n, 0.0);
  for (int j = 0; j < n; ++j)
    for (int k = 0; k < n; ++k) {
      const double br = B[2 * ((size_t)j * n + k)], bi = B[2 * ((size_t)j * n + k) + 1];
      for (int i = 0; i < n; ++i) {
        double ar, ai;
        if (!adjA) {
          ar = A[2 * ((size_t)k * n + i)];
          ai = A[2 * ((size_t)k * n + i) + 1];
        } else {
          ar = A[2 * ((size_t)i * n + k)];
          ai = -A[2 * ((size_t)i * n + k) + 1];
        }
        C[2 * ((size_t)j * n + i)] += ar * br - ai * bi;
        C[2 * ((size_t)j * n + i) + 1] += ar * bi + ai * br;
      }
    }
  return C;
}

int chain_create(ChainMem& cm, const dqmc_params* p, const dqmc_operators* o) {
  int rc = validate(p, o);
  if (rc) return rc;
  cm.p = *p;
  cm.L = p->L;
  cm.N = p->L * p->L;
  cm.n = 4 * cm.N;
  cm.M = p->slices;
  cm.nelem = p->slices / p->safe_mult + 1;
  const int n = cm.n, N = cm.N;
  const size_t nn = (size_t)n * n;
  ChainDesc& d = cm.d;
  memset(&d, 0, sizeof(d));
  d.L = cm.L;
  d.N = N;
  d.n = n;
  d.M = cm.M;
  d.safe_mult = p->safe_mult;
  d.nelem = cm.nelem;
  d.chkr = p->checkerboard;
  d.edrun = p->edrun;
  d.dtau = p->delta_tau;
  d.lam = p->lambda;
  d.r = p->r;
  d.c = p->c;
  d.u = p->u;
  void* ptr;
#define ALLOC(field, type, count)                                  \
  do {                                                             \
    if ((rc = alloc(cm, &ptr, sizeof(type) * (size_t)(count)))) return rc; \
    field = (type*)ptr;                                            \
  } while (0)
  for (int m = 0; m < kNumMat; ++m) {
    if ((m == M_ET2 || m == M_ETINV2 || m == M_ET || m == M_ETINV) && p->checkerboard) continue;
    ALLOC(d.mat[m], cplx, nn);
  }
  for (int v = 0; v < kNumVec; ++v) ALLOC(d.vec[v], double, n);
  ALLOC(d.ustack, cplx, nn * cm.nelem);
  ALLOC(d.tstack, cplx, nn * cm.nelem);
  ALLOC(d.dstack, double, (size_t)n * cm.nelem);
  ALLOC(d.phi, double, (size_t)3 * N * cm.M);
  ALLOC(d.qr_vn, double, 2 * n);
  ALLOC(d.qr_cand, cplx, (size_t)2 * kMaxWG * n);
  ALLOC(d.qr_cand_norm, double, 2 * kMaxWG);
  ALLOC(d.qr_cand_pos, int, 2 * kMaxWG);
  ALLOC(d.qr_perm, int, n);
  ALLOC(d.qr_lpos, int, n);
  ALLOC(d.qr_tau, cplx, n);
  ALLOC(d.qr_Tf, cplx, (size_t)16 * (n + 16));  // ceil(n / nb) panels of nb x nb, nb <= 16
  ALLOC(d.inv_piv, int, n);
  ALLOC(d.inv_bcast, cplx, 2 * n);
  ALLOC(d.lu_small, cplx, 32);
  ALLOC(d.lu_win, cplx, (size_t)2 * 4 * 4 * n);
  ALLOC(d.lu_diag, cplx, (size_t)2 * 16 * N);
  ALLOC(d.ld_A, cplx, (size_t)16 * 4 * n);
  ALLOC(d.ld_B, cplx, (size_t)16 * 4 * n);
  ALLOC(d.ld_cmd, int, 2);
  ALLOC(d.ucursor, long long, 1);
  ALLOC(d.scal, double, kNumScalar);
  ALLOC(d.log_count, long long, 1);
  unsigned int* bar;
  ALLOC(bar, unsigned int, 2);
  d.bar_count = bar;
  d.bar_gen = bar + 1;
  ALLOC(d.error, int, 1);
  HIPCHK(hipMemset(bar, 0, 2 * sizeof(unsigned int)));
  HIPCHK(hipMemset(d.error, 0, sizeof(int)));
  HIPCHK(hipMemset(d.ucursor, 0, sizeof(long long)));
  HIPCHK(hipMemset(d.log_count, 0, sizeof(long long)));
  HIPCHK(hipMemset(d.phi, 0, sizeof(double) * 3 * N * cm.M));
  {
    std::vector<double> sc(kNumScalar, 0.0);
    sc[S_MINRATIO] = 1e300;
    HIPCHK(hipMemcpy(d.scal, sc.data(), sizeof(double) * kNumScalar, hipMemcpyHostToDevice));
  }
  // operators
  int* nbd;
  ALLOC(nbd, int, 4 * N);
  HIPCHK(hipMemcpy(nbd, o->neighbors, sizeof(int) * 4 * N, hipMemcpyHostToDevice));
  d.nb = nbd;
  if (p->checkerboard == 2) {
    const int G = o->n_groups, nbonds = o->group_offsets[G];
    int *go, *bs;
    cplx* bb;
    double *em, *emi;
    ALLOC(go, int, G + 1);
    ALLOC(bs, int, 2 * nbonds);
    ALLOC(bb, cplx, (size_t)4 * nbonds * 16);
    ALLOC(em, double, n);
    ALLOC(emi, double, n);
    HIPCHK(hipMemcpy(go, o->group_offsets, sizeof(int) * (G + 1), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(bs, o->bond_sites, sizeof(int) * 2 * nbonds, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(bb, o->bond_blocks, sizeof(cplx) * 4 * nbonds * 16, hipMemcpyHostToDevice));
    std::vector<double> inv(n);
    for (int i = 0; i < n; ++i) inv[i] = 1.0 / o->expmu[i];
    HIPCHK(hipMemcpy(em, o->expmu, sizeof(double) * n, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(emi, inv.data(), sizeof(double) * n, hipMemcpyHostToDevice));
    d.ngroups = G;
    d.nbonds = nbonds;
    d.grp_off = go;
    d.bond_sites = bs;
    d.bond_blocks = bb;
    d.expmu = em;
    d.expmu_inv = emi;
    d.plaq_uniform = 0;
  } else if (p->checkerboard) {
    int* ps;
    cplx* pl;
    double *em, *emi;
    ALLOC(ps, int, 2 * N);
    ALLOC(pl, cplx, (size_t)4 * 2 * 4 * (N / 4) * 16);
    ALLOC(em, double, n);
    ALLOC(emi, double, n);
    HIPCHK(hipMemcpy(ps, o->plaq_sites, sizeof(int) * 2 * N, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(pl, o->plaq, sizeof(cplx) * 4 * 2 * 4 * (N / 4) * 16, hipMemcpyHostToDevice));
    std::vector<double> inv(n);
    for (int i = 0; i < n; ++i) inv[i] = 1.0 / o->expmu[i];
    HIPCHK(hipMemcpy(em, o->expmu, sizeof(double) * n, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(emi, inv.data(), sizeof(double) * n, hipMemcpyHostToDevice));
    d.plaq_sites = ps;
    d.plaq = pl;
    d.expmu = em;
    d.expmu_inv = emi;
    // without a magnetic field every plaquette of a (variant, group, block) carries the same
    // 4x4 block; the kernel then stages just those blocks in LDS
    {
      const int np = N / 4;
      const double* P = o->plaq;
      bool uni = true;
      for (int vgb = 0; vgb < 4 * 2 * 4 && uni; ++vgb)
        for (int q = 1; q < np && uni; ++q)
          uni = memcmp(P + 32 * ((size_t)vgb * np + q), P + 32 * ((size_t)vgb * np), 32 * sizeof(double)) == 0;
      d.plaq_uniform = uni ? 1 : 0;
    }
  } else {
    // B = eT (eT eV): the product of the two half-step exponentials is formed once
    std::vector<double> e2 = host_cmatmul(o->hop_exp, o->hop_exp, n, false);
    std::vector<double> ei2 = host_cmatmul(o->hop_exp_inv, o->hop_exp_inv, n, false);
    HIPCHK(hipMemcpy(d.mat[M_ET2], e2.data(), sizeof(cplx) * nn, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(d.mat[M_ETINV2], ei2.data(), sizeof(cplx) * nn, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(d.mat[M_ET], o->hop_exp, sizeof(cplx) * nn, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(d.mat[M_ETINV], o->hop_exp_inv, sizeof(cplx) * nn, hipMemcpyHostToDevice));
  }
#undef ALLOC
  return 0;
}

void chain_free(ChainMem& cm) {
  for (void* p : cm.allocs) (void)hipFree(p);
  cm.allocs.clear();
  if (cm.uniforms) (void)hipFree(cm.uniforms);
  if (cm.log_buf) (void)hipFree(cm.log_buf);
  for (void* p : {(void*)cm.grad, (void*)cm.grad_ptr, (void*)cm.mom, (void*)cm.phi_bak, (void*)cm.tdgf_mat,
                  (void*)cm.tdgf_vec, (void*)cm.gb, (void*)cm.meas})
    if (p) (void)hipFree(p);
  cm.meas = cm.meas_etgf = cm.meas_tnb = cm.meas_part = cm.meas_res = nullptr;
  cm.has_etgf = cm.has_hopping = false;
  cm.tdgf_mat = nullptr;
  cm.tdgf_vec = nullptr;
  cm.gb = nullptr;
  cm.uniforms = nullptr;
  cm.log_buf = nullptr;
  cm.grad = nullptr;
  cm.grad_ptr = nullptr;
  cm.mom = cm.phi_bak = nullptr;
}

int upload_desc(dqmc_batch* b, Group& g) {
  std::vector<ChainDesc> h;
  for (int ci : g.chains) h.push_back(b->chains[ci].d);
  if (!g.d_desc) HIPCHK(hipMalloc(&g.d_desc, sizeof(ChainDesc) * h.size()));
  HIPCHK(hipMemcpy(g.d_desc, h.data(), sizeof(ChainDesc) * h.size(), hipMemcpyHostToDevice));
  return 0;
}

// Workgroups per chain. The cooperative kernels of different shape groups run concurrently
// on their own streams, and a cooperative launch waits for all of its workgroups to be
// resident, so the budget of 256 (one workgroup per CU) is shared by every chain of the
// batch, not only by the chains of this group.
int default_groups(Group& g, int batch_chains) {
  int per = std::max(1, 256 / std::max(1, batch_chains));
  g.wg_qr = std::max(1, std::min({per, g.n / 8, kMaxWG}));
  g.wg_inv = std::max(std::max(1, (g.n + 511) / 512), std::min({per, g.n / 8, kMaxWG}));
  // local updates: the windowed variant is fastest with many narrow column blocks (the
  // own columns then fit in LDS beside the window); fall back to the n/8 blocks the other
  // variants were tuned at when the window does not fit
  g.wg_local = std::max(1, std::min({per, g.n / 4, kMaxWG}));
  if (local_window_sites(g.n, g.wg_local) < 2) g.wg_local = std::max(1, std::min({per, g.n / 8, kMaxWG}));
  // 8-column panels: each panel's serial part is shorter and the panel stays in registers
  // (the autotuner measures 16 too)
  g.qr_nb = 8;
  return 0;
}

// ---------------------------------------------------------------------------
// Group operations (enqueue only)
// ---------------------------------------------------------------------------
#define LCHK(expr)                                                                             \
  do {                                                                                         \
    hipError_t _e = (expr);                                                                    \
    if (_e != hipSuccess) return fail(DQMC_E_HIP, std::string(#expr) + ": " + hipGetErrorString(_e)); \
  } while (0)

// Per-class kernel timing with HIP events on the group's stream (dqmc_profile).
struct ProfScope {
  Group& g;
  int cls;
  hipEvent_t e0 = nullptr;
  ProfScope(Group& gr, int c) : g(gr), cls(c) {
    if (g.prof_mask & (1u << cls)) {
      e0 = g.take_event();
      (void)hipEventRecord(e0, g.stream);
    }
  }
  ~ProfScope() {
    if (e0) {
      hipEvent_t e1 = g.take_event();
      (void)hipEventRecord(e1, g.stream);
      g.prof.push_back({cls, e0, e1});
    }
  }
};

struct Ops {
  Group& g;
  explicit Ops(Group& gr) : g(gr) {}
  hipStream_t s() const { return g.stream; }
  const ChainDesc* d() const { return g.d_desc; }
  int nc() const { return g.nch(); }

  int copy(MatRef dst, MatRef src) {
    ProfScope ps(g, DQMC_K_OTHER);
    LCHK(launch_copy(s(), d(), nc(), g.n, dst, src));
    return 0;
  }
  int copyv(VecRef dst, VecRef src) {
    ProfScope ps(g, DQMC_K_OTHER);
    LCHK(launch_copy_vec(s(), d(), nc(), g.n, dst, src));
    return 0;
  }
  int gemm(MatRef C, MatRef A, GemmOp oa, MatRef B, GemmOp ob, VecRef dl = NOVEC, VecRef dr = NOVEC) {
    ProfScope ps(g, DQMC_K_GEMM);
    LCHK(launch_gemm(s(), d(), nc(), g.n, C, A, oa, B, ob, dl, dr, 0.0));
    return 0;
  }
  // decompose_udt! (linalg.jl:20-39). stack = true marks the column-graded products of the
  // stack ((B...B U) diag(D), add_slice_sequence_*), which take the blocked pre-pivoted route
  // unless udt route 1 is set; every other caller gets column-pivoted QR.
  int udt(MatRef A, VecRef D, MatRef Q, MatRef T, bool stack = false) {
    {
      ProfScope ps(g, DQMC_K_QRP);
      if (stack && g.udt_route == 0)
        LCHK(launch_qr_blocked(s(), d(), nc(), g.n, A, D, 1, g.wg_qr, g.qr_nb, g.qr_map, g.qr_map_blocks));
      else
        LCHK(launch_qrp(s(), d(), nc(), g.n, A, D, g.wg_qr));
    }
    ProfScope ps(g, DQMC_K_QFORM);
    if (stack && g.udt_route == 0) {
      LCHK(launch_form_T(s(), d(), nc(), g.n, A, D, T));
      LCHK(launch_form_Q_blocked(s(), d(), nc(), g.n, A, Q, g.qr_nb));
    } else {
      LCHK(launch_form_TQ(s(), d(), nc(), g.n, A, D, Q, T));
    }
    return 0;
  }
  int inverse(MatRef X) {
    ProfScope ps(g, DQMC_K_INVERSE);
    LCHK(launch_inverse(s(), d(), nc(), g.n, X, g.wg_inv));
    return 0;
  }
  // X <- op(B_slice) X / X op(B_slice); slice 0-based
  int applyB(MatRef X, BOp op, int slice) {
    if (g.chkr) {
      ProfScope ps(g, DQMC_K_APPLY_B);
      LCHK(launch_apply_B_chkr(s(), d(), nc(), g.n, g.N, X, op, slice, nullptr, g.ngroups));
      return 0;
    }
    // dense: B = eT2 eV, B^-1 = eV^-1 eTinv2, B^dagger = eV eT2^dagger
    const MatRef T4 = NM(M_TMP4);
    switch (op) {
      case B_LEFT:
        LCHK(launch_apply_eV(s(), d(), nc(), g.n, g.N, X, false, 0, 1.0, slice, nullptr));
        LCHK(launch_gemm(s(), d(), nc(), g.n, T4, NM(M_ET2), OP_N, X, OP_N, NOVEC, NOVEC, 0.0));
        return copy(X, T4);
      case B_RIGHT:
        LCHK(launch_gemm(s(), d(), nc(), g.n, T4, X, OP_N, NM(M_ET2), OP_N, NOVEC, NOVEC, 0.0));
        LCHK(launch_apply_eV(s(), d(), nc(), g.n, g.N, T4, true, 1, 1.0, slice, nullptr));
        return copy(X, T4);
      case B_INV_LEFT:
        LCHK(launch_gemm(s(), d(), nc(), g.n, T4, NM(M_ETINV2), OP_N, X, OP_N, NOVEC, NOVEC, 0.0));
        LCHK(launch_apply_eV(s(), d(), nc(), g.n, g.N, T4, false, 0, -1.0, slice, nullptr));
        return copy(X, T4);
      case B_INV_RIGHT:
        LCHK(launch_apply_eV(s(), d(), nc(), g.n, g.N, X, true, 1, -1.0, slice, nullptr));
        LCHK(launch_gemm(s(), d(), nc(), g.n, T4, X, OP_N, NM(M_ETINV2), OP_N, NOVEC, NOVEC, 0.0));
        return copy(X, T4);
      case B_DAG_LEFT:
        LCHK(launch_gemm(s(), d(), nc(), g.n, T4, NM(M_ET2), OP_C, X, OP_N, NOVEC, NOVEC, 0.0));
        LCHK(launch_apply_eV(s(), d(), nc(), g.n, g.N, T4, false, 2, 1.0, slice, nullptr));
        return copy(X, T4);
    }
    return fail(DQMC_E_ARG, "bad B op");
  }

  // wrap_greens! (stack.jl:316-325), slice 1-based as in the reference
  int wrap(MatRef X, int slice, int direction) {
    int rc;
    if (direction == -1) {
      if ((rc = applyB(X, B_INV_LEFT, slice - 2))) return rc;
      return applyB(X, B_RIGHT, slice - 2);
    }
    if ((rc = applyB(X, B_LEFT, slice - 1))) return rc;
    return applyB(X, B_INV_RIGHT, slice - 1);
  }

  int identity_elem(int k) {
    int rc;
    LCHK(launch_identity(s(), d(), nc(), g.n, US(k), NOVEC));
    LCHK(launch_identity(s(), d(), nc(), g.n, TS(k), NOVEC));
    LCHK(launch_ones_vec(s(), d(), nc(), g.n, DS(k)));
    (void)rc;
    return 0;
  }

  // build_stack (stack.jl:251-272)
  int build_stack() {
    int rc;
    if ((rc = identity_elem(0))) return rc;
    for (int i = 0; i < g.nelem - 1; ++i)
      if ((rc = add_left(i))) return rc;
    g.cur = g.M + 1;
    g.dir = -1;
    return 0;
  }

  // add_slice_sequence_left (stack.jl:278-293): element idx+1 from element idx (0-based)
  int add_left(int idx) {
    int rc;
    if ((rc = copy(NM(M_CURU), US(idx)))) return rc;
    for (int sl = idx * g.s; sl < (idx + 1) * g.s; ++sl)
      if ((rc = applyB(NM(M_CURU), B_LEFT, sl))) return rc;
    LCHK(launch_scale_cols(s(), d(), nc(), g.n, NM(M_CURU), DS(idx)));
    if ((rc = udt(NM(M_CURU), DS(idx + 1), US(idx + 1), NM(M_QRT), true))) return rc;
    return gemm(TS(idx + 1), NM(M_QRT), OP_N, TS(idx), OP_N);
  }

  // add_slice_sequence_right (stack.jl:299-313): element idx from element idx+1
  int add_right(int idx) {
    int rc;
    if ((rc = copy(NM(M_CURU), US(idx + 1)))) return rc;
    for (int sl = (idx + 1) * g.s - 1; sl >= idx * g.s; --sl)
      if ((rc = applyB(NM(M_CURU), B_DAG_LEFT, sl))) return rc;
    LCHK(launch_scale_cols(s(), d(), nc(), g.n, NM(M_CURU), DS(idx + 1)));
    if ((rc = udt(NM(M_CURU), DS(idx), US(idx), NM(M_QRT), true))) return rc;
    return gemm(TS(idx), NM(M_QRT), OP_N, TS(idx + 1), OP_N);
  }

  int load_left(int k) {
    int rc;
    if ((rc = copy(NM(M_UL), US(k)))) return rc;
    if ((rc = copy(NM(M_TL), TS(k)))) return rc;
    return copyv(NV(V_DL), DS(k));
  }
  int load_right(int k) {
    int rc;
    if ((rc = copy(NM(M_UR), US(k)))) return rc;
    if ((rc = copy(NM(M_TR), TS(k)))) return rc;
    return copyv(NV(V_DR), DS(k));
  }

  // calculate_greens (stack.jl:338-369) [+ calculate_logdet (stack.jl:377-385)]
  int calculate_greens(bool logdet) {
    return g.greens_route == 0 ? calculate_greens_scalesep() : calculate_greens_reference(logdet);
  }

  // G = [1 + Ul Dl Tl (Ur Dr Tr)^H]^-1 with the D scales separated into Db = max(D, 1)
  // and Ds = min(D, 1):
  //   G = Ur Drb^-1 [Dlb^-1 Ul^H Ur Drb^-1 + Dls Tl Tr^H Drs]^-1 Dlb^-1 Ul^H,
  // the bracket being well conditioned, so one unpivoted Householder QR and a solve replace
  // the two inner UDT decompositions of the verbatim route; log|det G| = -log|det [.]|
  // - sum log Dlb - sum log Drb (Ul, Ur unitary). Always leaves log|det G| in S_LOGDET.
  int calculate_greens_scalesep() {
    int rc;
    LCHK(launch_split_scales(s(), d(), nc(), g.n, NV(V_DL), NV(V_D1), NV(V_D2)));  // 1/Dlb, Dls
    LCHK(launch_split_scales(s(), d(), nc(), g.n, NV(V_DR), NV(V_D3), NV(V_D4)));  // 1/Drb, Drs
    {
      ProfScope ps(g, DQMC_K_GEMM);
      LCHK(launch_gemm(s(), d(), nc(), g.n, NM(M_TMP), NM(M_UL), OP_C, NM(M_UR), OP_N, NV(V_D1), NV(V_D3), 0.0));
      LCHK(launch_gemm(s(), d(), nc(), g.n, NM(M_TMP), NM(M_TL), OP_N, NM(M_TR), OP_C, NV(V_D2), NV(V_D4), 1.0));
    }
    {
      // [.] = Q R (Householder, no pivoting: the matrix is well conditioned), then
      // Y = [.]^-1 Dlb^-1 Ul^H = R^-1 Q^H (Dlb^-1 Ul^H); log|det [.]| = sum log |R_kk|
      ProfScope ps(g, DQMC_K_INVERSE);
      LCHK(launch_qr_blocked(s(), d(), nc(), g.n, NM(M_TMP), NV(V_D5), 0, g.wg_qr, g.qr_nb, g.qr_map,
                             g.qr_map_blocks));
      LCHK(launch_adjoint(s(), d(), nc(), g.n, NM(M_TMP2), NM(M_UL)));
      LCHK(launch_scale_rows_copy(s(), d(), nc(), g.n, NM(M_TMP2), NM(M_TMP2), NV(V_D1)));
      LCHK(launch_qr_solve(s(), d(), nc(), g.n, NM(M_TMP), NM(M_TMP2), g.qr_nb));
      LCHK(launch_scale_rows_copy(s(), d(), nc(), g.n, NM(M_TMP2), NM(M_TMP2), NV(V_D3)));
      LCHK(launch_sum_log(s(), d(), nc(), g.n, NV(V_D5), S_LOGDET_A, 0));
    }
    if ((rc = gemm(NM(M_G), NM(M_UR), OP_N, NM(M_TMP2), OP_N))) return rc;
    LCHK(launch_sum_log(s(), d(), nc(), g.n, NV(V_D1), S_LOGDET, 0));
    LCHK(launch_sum_log(s(), d(), nc(), g.n, NV(V_D3), S_LOGDET, 1));
    LCHK(launch_negate_scalar(s(), d(), nc(), S_LOGDET_A));
    LCHK(launch_add_scalar(s(), d(), nc(), S_LOGDET, S_LOGDET_A));
    return 0;
  }

  // calculate_greens (stack.jl:338-369) verbatim: two inner UDT decompositions.
  int calculate_greens_reference(bool logdet) {
    int rc;
    // tmp = diag(Dl) Tl Tr^H diag(Dr)
    if ((rc = gemm(NM(M_TMP), NM(M_TL), OP_N, NM(M_TR), OP_C, NV(V_DL), NV(V_DR)))) return rc;
    // U1, D1, T1 = udt(tmp)
    if ((rc = udt(NM(M_TMP), NV(V_D1), NM(M_QRQ), NM(M_QRT)))) return rc;
    // s.U = Ul U1  -> TMP2
    if ((rc = gemm(NM(M_TMP2), NM(M_UL), OP_N, NM(M_QRQ), OP_N))) return rc;
    // tmp2 = T1 Ur^H -> TMP3
    if ((rc = gemm(NM(M_TMP3), NM(M_QRT), OP_N, NM(M_UR), OP_C))) return rc;
    // tmp = s.U^H / tmp2 (myrdiv, linalg.jl:61) = s.U^H inv(tmp2)
    if ((rc = copy(NM(M_TMP4), NM(M_TMP3)))) return rc;
    if ((rc = inverse(NM(M_TMP4)))) return rc;
    if ((rc = gemm(NM(M_TMP), NM(M_TMP2), OP_C, NM(M_TMP4), OP_N))) return rc;
    // tmp += diag(D1); u, d, t = udt(tmp)
    LCHK(launch_add_diag(s(), d(), nc(), g.n, NM(M_TMP), NV(V_D1)));
    if ((rc = udt(NM(M_TMP), NV(V_D2), NM(M_QRQ), NM(M_QRT)))) return rc;
    // s.T = inv(t tmp2)
    if ((rc = gemm(NM(M_TG), NM(M_QRT), OP_N, NM(M_TMP3), OP_N))) return rc;
    if ((rc = inverse(NM(M_TG)))) return rc;
    if (logdet) LCHK(launch_sum_log(s(), d(), nc(), g.n, NV(V_D2), S_LOGDET, 0));  // log d (before 1/d)
    // s.U = (s.U u)^H
    if ((rc = gemm(NM(M_TMP), NM(M_TMP2), OP_N, NM(M_QRQ), OP_N))) return rc;
    LCHK(launch_adjoint(s(), d(), nc(), g.n, NM(M_UG), NM(M_TMP)));
    // s.d = 1 ./ s.d
    LCHK(launch_recip_vec(s(), d(), nc(), g.n, NV(V_DG), NV(V_D2)));
    // greens = s.T diag(s.d) s.U
    LCHK(launch_scale_rows_copy(s(), d(), nc(), g.n, NM(M_TMP3), NM(M_UG), NV(V_DG)));
    if ((rc = gemm(NM(M_G), NM(M_TG), OP_N, NM(M_TMP3), OP_N))) return rc;
    if (logdet) {
      // log_det = log|det U| + sum log(1/d) + log|det T|, with log|det T| = -log|det(t tmp2)|
      // (the inverse above left log|det(t tmp2)| in S_LOGDET_A) and sum log(1/d) = -sum log d
      LCHK(launch_add_scalar(s(), d(), nc(), S_LOGDET, S_LOGDET_A));  // S = log d + log|det(t tmp2)|
      LCHK(launch_negate_scalar(s(), d(), nc(), S_LOGDET));
      if ((rc = copy(NM(M_TMP4), NM(M_UG)))) return rc;
      if ((rc = inverse(NM(M_TMP4)))) return rc;  // log|det U| into S_LOGDET_A
      LCHK(launch_add_scalar(s(), d(), nc(), S_LOGDET, S_LOGDET_A));
    }
    return 0;
  }

  // propagate (stack.jl:391-499)
  int propagate() {
    int rc;
    const int M = g.M, sm = g.s;
    g.recomputed = false;
    if (g.dir == 1) {
      if (g.cur % sm == 0) {
        g.cur += 1;
        if (g.cur == 1) {
          if ((rc = load_right(0))) return rc;
          if ((rc = identity_elem(0))) return rc;
          if ((rc = load_left(0))) return rc;
          return calculate_greens(true);
        } else if (g.cur <= M) {
          const int idx = (g.cur - 1) / sm;
          if ((rc = load_right(idx))) return rc;
          if ((rc = add_left(idx - 1))) return rc;
          if ((rc = load_left(idx))) return rc;
          if (g.all_checks) {
            if ((rc = copy(NM(M_GW), NM(M_G)))) return rc;
            if ((rc = wrap(NM(M_GW), g.cur - 1, 1))) return rc;
          }
          if ((rc = calculate_greens(false))) return rc;
          if (g.all_checks) {
            LCHK(launch_maxdiff(s(), d(), nc(), g.n, NM(M_GW), NM(M_G)));
            g.recomputed = true;
          }
          return 0;
        } else {
          if ((rc = add_left(g.nelem - 2))) return rc;
          g.dir = -1;
          g.cur = M + 1;
          return propagate();
        }
      }
      if ((rc = wrap(NM(M_G), g.cur, 1))) return rc;
      g.cur += 1;
      return 0;
    }
    if ((g.cur - 1) % sm == 0) {
      g.cur -= 1;
      if (g.cur == M) {
        const int last = g.nelem - 1;
        if ((rc = load_left(last))) return rc;
        if ((rc = identity_elem(last))) return rc;
        if ((rc = load_right(last))) return rc;
        if ((rc = calculate_greens(true))) return rc;
        return wrap(NM(M_G), g.cur + 1, -1);
      } else if (g.cur > 0) {
        const int idx = g.cur / sm + 1;
        if ((rc = load_left(idx - 1))) return rc;
        if ((rc = add_right(idx - 1))) return rc;
        if ((rc = load_right(idx - 1))) return rc;
        if (g.all_checks)
          if ((rc = copy(NM(M_GW), NM(M_G)))) return rc;
        if ((rc = calculate_greens(false))) return rc;
        if (g.all_checks) {
          LCHK(launch_maxdiff(s(), d(), nc(), g.n, NM(M_GW), NM(M_G)));
          g.recomputed = true;
        }
        return wrap(NM(M_G), g.cur + 1, -1);
      } else {
        if ((rc = add_right(0))) return rc;
        g.dir = 1;
        g.cur = 0;
        return propagate();
      }
    }
    if ((rc = wrap(NM(M_G), g.cur, -1))) return rc;
    g.cur -= 1;
    return 0;
  }

  // calc_Bchain (op B_LEFT), calc_Bchain_inv (B_INV_LEFT), calc_Bchain_dagger (B_DAG_LEFT)
  // (fermion_measurements.jl:964-1058): U D T of the product over slices start..stop (1-based,
  // inclusive), decomposed at every multiple of safe_mult and at the last slice taken.
  int bchain(int start, int stop, BOp op, MatRef U, VecRef D, MatRef T) {
    int rc;
    LCHK(launch_identity(s(), d(), nc(), g.n, U, NOVEC));
    LCHK(launch_identity(s(), d(), nc(), g.n, T, NOVEC));
    LCHK(launch_ones_vec(s(), d(), nc(), g.n, D));
    const bool down = op != B_LEFT;
    for (int t = 0; t <= stop - start; ++t) {
      const int k = down ? stop - t : start + t;
      if ((rc = applyB(U, op, k - 1))) return rc;
      if (k % g.s == 0 || k == (down ? start : stop)) {
        LCHK(launch_scale_cols(s(), d(), nc(), g.n, U, D));
        if ((rc = udt(U, D, NM(M_QRQ), NM(M_QRT)))) return rc;
        if ((rc = copy(U, NM(M_QRQ)))) return rc;
        if ((rc = gemm(NM(M_TMP), NM(M_QRT), OP_N, T, OP_N))) return rc;
        if ((rc = copy(T, NM(M_TMP)))) return rc;
      }
    }
    return 0;
  }

  // ---- time-displaced Green's functions (fermion_measurements.jl:1239-1541) ----------------
  // Storage in ChainDesc::xmat, nr = M / safe_mult checkpoints per chain q (0 BT0Inv, 1 BBetaT,
  // 2 BT0, 3 BBetaTInv): U of entry k at 2 (q nr + k), T at 2 (q nr + k) + 1, D at xvec q nr + k;
  // then Gt0[tau] at 8 nr + tau - 1 and G0t[tau] at 8 nr + M + tau - 1 (tau = 1..M).
  int nr() const { return g.M / g.s; }
  MatRef XU(int q, int k) const { return MatRef{MK_XMAT, 2 * (q * nr() + k)}; }
  MatRef XT(int q, int k) const { return MatRef{MK_XMAT, 2 * (q * nr() + k) + 1}; }
  VecRef XD(int q, int k) const { return VecRef{VK_XVEC, q * nr() + k}; }
  MatRef GT0(int tau) const { return MatRef{MK_XMAT, 8 * nr() + tau - 1}; }
  MatRef G0T(int tau) const { return MatRef{MK_XMAT, 8 * nr() + g.M + tau - 1}; }

  // effective_greens2greens! (fermion_measurements.jl:1125-1141, 1178-1185):
  // X <- hop_half_inv[A] hop_half_inv[B] X hop_half[B] hop_half[A], or eT^-1 X eT. X != M_TMP4.
  int effective_greens(MatRef X) {
    if (g.chkr) {
      ProfScope ps(g, DQMC_K_APPLY_B);
      LCHK(launch_apply_B_chkr(s(), d(), nc(), g.n, g.N, X, B_EFF_RIGHT, 0, nullptr, g.ngroups));
      LCHK(launch_apply_B_chkr(s(), d(), nc(), g.n, g.N, X, B_EFF_INV_LEFT, 0, nullptr, g.ngroups));
      return 0;
    }
    int rc;
    if ((rc = gemm(NM(M_TMP4), X, OP_N, NM(M_ET), OP_N))) return rc;
    return gemm(X, NM(M_ETINV), OP_N, NM(M_TMP4), OP_N);
  }
  // the U, T form (fermion_measurements.jl:1143-1159, 1187-1193): T <- T hop_half[B] hop_half[A],
  // U <- hop_half_inv[A] hop_half_inv[B] U (eT / eT^-1 without the checkerboard)
  int effective_greens_ut(MatRef U, MatRef T) {
    if (g.chkr) {
      ProfScope ps(g, DQMC_K_APPLY_B);
      LCHK(launch_apply_B_chkr(s(), d(), nc(), g.n, g.N, T, B_EFF_RIGHT, 0, nullptr, g.ngroups));
      LCHK(launch_apply_B_chkr(s(), d(), nc(), g.n, g.N, U, B_EFF_INV_LEFT, 0, nullptr, g.ngroups));
      return 0;
    }
    int rc;
    if ((rc = gemm(NM(M_TMP4), T, OP_N, NM(M_ET), OP_N))) return rc;
    if ((rc = copy(T, NM(M_TMP4)))) return rc;
    if ((rc = gemm(NM(M_TMP4), NM(M_ETINV), OP_N, U, OP_N))) return rc;
    return copy(U, NM(M_TMP4));
  }

  // calc_Bchain_udts! (fermion_measurements.jl:1425-1507) into chain q. Entries are indexed as
  // after the reference's final reverse!: entry k covers slices 1 .. (k+1)s (left chains) or
  // k s + 1 .. M (right chains).
  int bchain_udts(int q, bool invert, bool right) {
    int rc;
    const bool rightmult = (right && !invert) || (!right && invert);
    const int R = nr();
    const BOp op = !invert ? (right ? B_RIGHT : B_LEFT) : (right ? B_INV_LEFT : B_INV_RIGHT);
    const MatRef cur = NM(M_CURU);
    for (int i = 0; i < R; ++i) {
      const int k = right ? R - 1 - i : i;   // entry written by this step
      const int kp = right ? k + 1 : k - 1;  // entry written by the previous step
      if (i == 0)
        LCHK(launch_identity(s(), d(), nc(), g.n, cur, NOVEC));
      else if ((rc = copy(cur, rightmult ? XT(q, kp) : XU(q, kp))))
        return rc;
      for (int j = 0; j < g.s; ++j) {
        const int sl = right ? (k + 1) * g.s - 1 - j : k * g.s + j;  // 0-based slice
        if ((rc = applyB(cur, op, sl))) return rc;
      }
      if (i != 0) {
        if (!rightmult)
          LCHK(launch_scale_cols(s(), d(), nc(), g.n, cur, XD(q, kp)));
        else
          LCHK(launch_scale_rows_copy(s(), d(), nc(), g.n, cur, cur, XD(q, kp)));
      }
      if (!rightmult) {
        if ((rc = udt(cur, XD(q, k), XU(q, k), i == 0 ? XT(q, k) : NM(M_QRT)))) return rc;
        if (i != 0 && (rc = gemm(XT(q, k), NM(M_QRT), OP_N, XT(q, kp), OP_N))) return rc;
      } else {
        if ((rc = udt(cur, XD(q, k), i == 0 ? XU(q, k) : NM(M_QRQ), XT(q, k)))) return rc;
        if (i != 0 && (rc = gemm(XU(q, k), XU(q, kp), OP_N, NM(M_QRQ), OP_N))) return rc;
      }
    }
    return 0;
  }

  // res = [1 + U D T]^-1 (inv_one_plus_udt_scalettar!, linalg.jl:310-339); T^-1 stays in M_TMP
  int inv_one_plus_udt_scalettar(MatRef res, MatRef U, VecRef D, MatRef T) {
    int rc;
    LCHK(launch_split_scales(s(), d(), nc(), g.n, D, NV(V_D1), NV(V_D2)));  // 1/Dp, Dm
    if ((rc = copy(NM(M_TMP), T))) return rc;
    if ((rc = inverse(NM(M_TMP)))) return rc;
    if ((rc = copy(NM(M_TMP2), NM(M_TMP)))) return rc;
    LCHK(launch_scale_cols(s(), d(), nc(), g.n, NM(M_TMP2), NV(V_D1)));            // l = T \ Dp^-1
    LCHK(launch_add_scaled_cols(s(), d(), nc(), g.n, NM(M_TMP2), U, NV(V_D2)));    // r = U Dm + l
    if ((rc = udt(NM(M_TMP2), NV(V_D3), NM(M_QRQ), NM(M_QRT)))) return rc;         // u, d, t
    if ((rc = inverse(NM(M_QRT)))) return rc;
    LCHK(launch_recip_vec(s(), d(), nc(), g.n, NV(V_D4), NV(V_D3)));
    LCHK(launch_scale_cols(s(), d(), nc(), g.n, NM(M_QRT), NV(V_D4)));             // r = t \ d^-1
    if ((rc = gemm(NM(M_TMP3), NM(M_QRT), OP_N, NM(M_QRQ), OP_C, NV(V_D1)))) return rc;  // Dp^-1 r u'
    if ((rc = udt(NM(M_TMP3), NV(V_D3), NM(M_QRQ), NM(M_QRT)))) return rc;
    if ((rc = gemm(NM(M_TMP2), NM(M_TMP), OP_N, NM(M_QRQ), OP_N, NOVEC, NV(V_D3)))) return rc;  // T \ u d
    return gemm(res, NM(M_TMP2), OP_N, NM(M_QRT), OP_N);
  }

  // res = [Ua Da Ta + Ub Db Tb]^-1 (inv_sum_udts_scalettar!, linalg.jl:512-567 and its
  // allocating form :442-496); Ta Tb^-1 and Tb \ U use one explicit Tb^-1.
  int inv_sum_udts_scalettar(MatRef res, MatRef Ua, VecRef Da, MatRef Ta, MatRef Ub, VecRef Db, MatRef Tb,
                             MatRef Uout = MatRef{-1, 0}, VecRef Dout = NOVEC, MatRef Tout = MatRef{-1, 0}) {
    int rc;
    LCHK(launch_split_scales(s(), d(), nc(), g.n, Da, NV(V_D1), NV(V_D2)));  // 1/Dap, Dam
    LCHK(launch_split_scales(s(), d(), nc(), g.n, Db, NV(V_D3), NV(V_D4)));  // 1/Dbp, Dbm
    if ((rc = copy(NM(M_TMP), Tb))) return rc;
    if ((rc = inverse(NM(M_TMP)))) return rc;  // Tb^-1
    if ((rc = gemm(NM(M_TMP2), Ta, OP_N, NM(M_TMP), OP_N, NV(V_D2), NV(V_D3)))) return rc;
    {
      ProfScope ps(g, DQMC_K_GEMM);
      LCHK(launch_gemm(s(), d(), nc(), g.n, NM(M_TMP2), Ua, OP_C, Ub, OP_N, NV(V_D1), NV(V_D4), 1.0));
    }
    if ((rc = udt(NM(M_TMP2), NV(V_D5), NM(M_QRQ), NM(M_QRT)))) return rc;
    LCHK(launch_scale_rows_copy(s(), d(), nc(), g.n, NM(M_QRT), NM(M_QRT), NV(V_D5)));  // D T
    if ((rc = inverse(NM(M_QRT)))) return rc;
    if ((rc = gemm(NM(M_TMP3), NM(M_QRT), OP_N, NM(M_QRQ), OP_C, NV(V_D3), NV(V_D1)))) return rc;
    if ((rc = udt(NM(M_TMP3), NV(V_D5), NM(M_QRQ), NM(M_QRT)))) return rc;
    if (Uout.kind >= 0) {
      // U, D, T form: U = Tb^-1 U, T = T Ua'
      if ((rc = gemm(Uout, NM(M_TMP), OP_N, NM(M_QRQ), OP_N))) return rc;
      if ((rc = gemm(Tout, NM(M_QRT), OP_N, Ua, OP_C))) return rc;
      return copyv(Dout, NV(V_D5));
    }
    if ((rc = gemm(NM(M_TMP2), NM(M_TMP), OP_N, NM(M_QRQ), OP_N, NOVEC, NV(V_D5)))) return rc;
    if ((rc = gemm(NM(M_TMP3), NM(M_QRT), OP_N, Ua, OP_C))) return rc;
    return gemm(res, NM(M_TMP2), OP_N, NM(M_TMP3), OP_N);
  }

  // measure_tdgfs! (fermion_measurements.jl:1343-1407) with fill_tdgf! (:1513-1541).
  // G0t is negated at the checkpoint slices instead of after the fill; the fill is linear and
  // IEEE negation commutes exactly with it, so the values are the same.
  int measure_tdgfs() {
    int rc;
    if ((rc = bchain_udts(0, true, false))) return rc;   // BT0Inv
    if ((rc = bchain_udts(1, false, true))) return rc;   // BBetaT
    if ((rc = bchain_udts(2, false, false))) return rc;  // BT0
    if ((rc = bchain_udts(3, true, true))) return rc;    // BBetaTInv
    const int R = nr(), M = g.M, sm = g.s;
    for (int i = 0; i < R; ++i) {
      const int tau = i * sm + 1;
      if (i != 0) {
        if ((rc = inv_sum_udts_scalettar(GT0(tau), XU(0, i - 1), XD(0, i - 1), XT(0, i - 1), XU(1, i), XD(1, i),
                                         XT(1, i))))
          return rc;
        if ((rc = inv_sum_udts_scalettar(G0T(tau), XU(2, i - 1), XD(2, i - 1), XT(2, i - 1), XU(3, i), XD(3, i),
                                         XT(3, i))))
          return rc;
      } else {
        if ((rc = inv_one_plus_udt_scalettar(GT0(1), XU(1, 0), XD(1, 0), XT(1, 0)))) return rc;
        if ((rc = inv_one_plus_udt_scalettar(G0T(1), XU(3, 0), XD(3, 0), XT(3, 0)))) return rc;
      }
      if ((rc = effective_greens(GT0(tau)))) return rc;
      if ((rc = effective_greens(G0T(tau)))) return rc;
      LCHK(launch_negate(s(), d(), nc(), g.n, G0T(tau)));
    }
    const int mhalf = M / 2 + 1;
    for (int tau = mhalf; tau <= M; ++tau) {
      if ((tau - 1) % sm == 0) continue;
      if ((rc = copy(GT0(tau), GT0(tau - 1)))) return rc;
      if ((rc = applyB(GT0(tau), B_LEFT, tau - 1))) return rc;
      if ((rc = copy(G0T(tau), G0T(tau - 1)))) return rc;
      if ((rc = applyB(G0T(tau), B_INV_RIGHT, tau - 1))) return rc;
    }
    for (int tau = mhalf - 1; tau >= 1; --tau) {
      if ((tau - 1) % sm == 0) continue;
      if ((rc = copy(GT0(tau), GT0(tau + 1)))) return rc;
      if ((rc = applyB(GT0(tau), B_INV_LEFT, tau))) return rc;  // B(tau + 1)^-1
      if ((rc = copy(G0T(tau), G0T(tau + 1)))) return rc;
      if ((rc = applyB(G0T(tau), B_RIGHT, tau))) return rc;
    }
    return 0;
  }

  int local_updates(double box, int dolog) {
    if (g.cur < 1 || g.cur > g.M) return fail(DQMC_E_STATE, "local_updates outside slices 1..M");
    ProfScope ps(g, DQMC_K_LOCAL);
    const bool tile_fits = local_tile_fits(g.n, g.wg_local), delay_fits = local_delayed_fits(g.n);
    if (g.local_variant == 2 && !tile_fits)
      return fail(DQMC_E_ARG, "G tiles do not fit in LDS at this workgroup count");
    if (g.local_variant == 3 && !delay_fits) return fail(DQMC_E_ARG, "delayed updates do not fit in LDS at this size");
    if (g.local_variant == 4 && local_window_sites(g.n, g.wg_local) < 2)
      return fail(DQMC_E_ARG, "windowed updates do not fit in LDS at this size and workgroup count");
    int variant = g.local_variant;
    if (variant == 0) variant = local_window_sites(g.n, g.wg_local) >= 2 ? 4 : tile_fits ? 2 : delay_fits ? 3 : 1;
    LCHK(launch_local_updates(s(), d(), nc(), g.n, g.cur - 1, nullptr, box, g.wg_local, dolog, variant));
    return 0;
  }
};

int check_device_error(dqmc_batch* b) {
  for (auto& cm : b->chains) {
    int e = 0;
    HIPCHK(hipMemcpy(&e, cm.d.error, sizeof(int), hipMemcpyDeviceToHost));
    if (e == 1) return fail(DQMC_E_DEVICE, "device: chain-group barrier timed out");
    if (e == 2) return fail(DQMC_E_DEVICE, "device: uniform stream exhausted (upload more with set_uniforms)");
  }
  return 0;
}

int batch_sync(dqmc_batch* b) {
  for (auto& g : b->groups) HIPCHK(hipStreamSynchronize(g.stream));
  if (b->mstream) HIPCHK(hipStreamSynchronize(b->mstream));
  return check_device_error(b);
}

int batch_create(const dqmc_params* p, const dqmc_operators* ops, int nchains, int device, dqmc_batch** out) {
  if (!out || nchains < 1 || !p || !ops) return fail(DQMC_E_ARG, "bad batch arguments");
  for (int i = 0; i < nchains; ++i) {
    int rc = validate(&p[i], &ops[i]);
    if (rc) return rc;
  }
  HIPCHK(hipSetDevice(device));
  auto* b = new dqmc_batch();
  b->device = device;
  b->chains.resize(nchains);
  for (int i = 0; i < nchains; ++i) {
    int rc = chain_create(b->chains[i], &p[i], &ops[i]);
    if (rc) {
      for (auto& cm : b->chains) chain_free(cm);
      delete b;
      return rc;
    }
  }
  // shape groups: identical (L, M, safe_mult, checkerboard, all_checks)
  std::map<std::tuple<int, int, int, int, int>, int> idx;
  b->chain_group.resize(nchains);
  for (int i = 0; i < nchains; ++i) {
    const dqmc_params& q = p[i];
    auto key = std::make_tuple(q.L, q.slices, q.safe_mult, q.checkerboard, q.all_checks);
    auto it = idx.find(key);
    int gi;
    if (it == idx.end()) {
      gi = (int)b->groups.size();
      idx[key] = gi;
      Group g;
      g.L = q.L;
      g.N = q.L * q.L;
      g.n = 4 * g.N;
      g.M = q.slices;
      g.s = q.safe_mult;
      g.nelem = q.slices / q.safe_mult + 1;
      g.chkr = q.checkerboard;
      g.ngroups = q.checkerboard == 2 ? ops[i].n_groups : 0;
      g.all_checks = q.all_checks != 0;
      b->groups.push_back(g);
    } else {
      gi = it->second;
    }
    b->groups[gi].chains.push_back(i);
    b->chain_group[i] = gi;
  }
  for (auto& g : b->groups) {
    HIPCHK(hipStreamCreateWithFlags(&g.stream, hipStreamNonBlocking));
    default_groups(g, nchains);
    int rc = upload_desc(b, g);
    if (rc) return rc;
  }
  if (b->groups.size() > 1) HIPCHK(hipStreamCreateWithFlags(&b->mstream, hipStreamNonBlocking));
  HIPCHK(hipMalloc(&b->d_scratch, sizeof(double) * 128));
  *out = b;
  return 0;
}

int batch_destroy(dqmc_batch* b) {
  if (!b) return 0;
  (void)hipSetDevice(b->device);
  for (auto& g : b->groups)
    if (g.stream) (void)hipStreamSynchronize(g.stream);
  for (auto& g : b->groups) {
    g.drop_graphs();
    // timing events: pending records and the reuse pool
    for (auto& r : g.prof) {
      (void)hipEventDestroy(r.a);
      (void)hipEventDestroy(r.b);
    }
    g.prof.clear();
    for (hipEvent_t e : g.event_pool) (void)hipEventDestroy(e);
    g.event_pool.clear();
    if (g.stream) (void)hipStreamDestroy(g.stream);
    if (g.d_desc) (void)hipFree(g.d_desc);
  }
  for (auto& cm : b->chains) chain_free(cm);
  if (b->d_scratch) (void)hipFree(b->d_scratch);
  b->drop_mgraphs();
  if (b->mstream) (void)hipStreamDestroy(b->mstream);
  if (b->d_mplan) (void)hipFree(b->d_mplan);
  if (b->d_mdesc) (void)hipFree(b->d_mdesc);
  delete b;
  return 0;
}

int set_uniforms(dqmc_batch* b, int chain, const double* u, long long count) {
  if (chain < 0 || chain >= (int)b->chains.size() || !u || count < 0) return fail(DQMC_E_ARG, "bad uniforms");
  ChainMem& cm = b->chains[chain];
  Group& g = b->groups[b->chain_group[chain]];
  HIPCHK(hipStreamSynchronize(g.stream));
  if (count > cm.ucap) {
    if (cm.uniforms) HIPCHK(hipFree(cm.uniforms));
    HIPCHK(hipMalloc(&cm.uniforms, sizeof(double) * std::max<long long>(count, 1)));
    cm.ucap = count;
  }
  HIPCHK(hipMemcpy(cm.uniforms, u, sizeof(double) * count, hipMemcpyHostToDevice));
  long long zero = 0;
  HIPCHK(hipMemcpy(cm.d.ucursor, &zero, sizeof(zero), hipMemcpyHostToDevice));
  cm.d.uniforms = cm.uniforms;
  cm.d.n_uniforms = count;
  return upload_desc(b, g);
}

int get_stats(dqmc_batch* b, int chain, dqmc_stats* st) {
  if (!st || chain < 0 || chain >= (int)b->chains.size()) return fail(DQMC_E_ARG, "bad stats args");
  int rc = batch_sync(b);
  ChainMem& cm = b->chains[chain];
  Group& g = b->groups[b->chain_group[chain]];
  double sc[kNumScalar];
  HIPCHK(hipMemcpy(sc, cm.d.scal, sizeof(sc), hipMemcpyDeviceToHost));
  long long cur = 0;
  HIPCHK(hipMemcpy(&cur, cm.d.ucursor, sizeof(cur), hipMemcpyDeviceToHost));
  int err = 0;
  HIPCHK(hipMemcpy(&err, cm.d.error, sizeof(err), hipMemcpyDeviceToHost));
  st->current_slice = g.cur;
  st->direction = g.dir;
  st->boson_action = sc[S_BOSON_ACTION];
  st->log_det = sc[S_LOGDET];
  st->drift_last = sc[S_DRIFT_LAST];
  st->drift_max = sc[S_DRIFT_MAX];
  st->acc_last = sc[S_ACC_SLICE] / cm.N;
  st->accepted = sc[S_ACC_TOTAL];
  st->proposed = sc[S_PROP_TOTAL];
  st->nonreal = sc[S_NONREAL];
  st->min_ratio_gap = sc[S_MINRATIO];
  st->uniforms_used = cur;
  st->device_error = err;
  st->global_accepted = sc[S_GLOBAL_NACC];
  st->global_proposed = sc[S_GLOBAL_NPROP];
  (void)rc;
  return 0;
}

int group_init(dqmc_batch*, Group& g) {
  Ops o(g);
  int rc;
  LCHK(launch_boson_action(g.stream, g.d_desc, g.nch(), g.N));
  LCHK(launch_ones_vec(g.stream, g.d_desc, g.nch(), g.n, NV(V_ONES)));
  if ((rc = o.build_stack())) return rc;
  return o.propagate();
}

int group_sweeps(Group& g, int nsweeps, double box) {
  Ops o(g);
  int rc;
  for (int sw = 0; sw < nsweeps; ++sw)
    for (int step = 0; step < g.M; ++step) {
      if ((rc = o.propagate())) return rc;
      if ((rc = o.local_updates(box, 0))) return rc;
    }
  return 0;
}

// One sweep of a group through a HIP graph. The work of a sweep is fixed by the group's
// (current_slice, direction) at its start, the box and the logging flag -- propagate is a
// host-side state machine and every kernel reads the chains' state from device memory -- so
// each starting state is captured once and replayed; the host state then jumps to where
// the captured sweep ends.
int group_sweep_graph(Group& g, double box, int dolog) {
  for (auto& e : g.graphs)
    if (e.cur0 == g.cur && e.dir0 == g.dir && e.dolog == dolog && e.box == box) {
      HIPCHK(hipGraphLaunch(e.exec, g.stream));
      g.cur = e.cur1;
      g.dir = e.dir1;
      return 0;
    }
  Group::SweepGraph e{g.cur, g.dir, dolog, box, 0, 0, nullptr};
  HIPCHK(hipStreamBeginCapture(g.stream, hipStreamCaptureModeRelaxed));
  int rc = 0;
  {
    Ops o(g);
    for (int step = 0; step < g.M && !rc; ++step) {
      rc = o.propagate();
      if (!rc) rc = o.local_updates(box, dolog);
    }
  }
  hipGraph_t graph = nullptr;
  const hipError_t ce = hipStreamEndCapture(g.stream, &graph);
  hipError_t ie = hipSuccess;
  if (!rc && ce == hipSuccess) ie = hipGraphInstantiate(&e.exec, graph, nullptr, nullptr, 0);
  if (graph) (void)hipGraphDestroy(graph);
  if (rc || ce != hipSuccess || ie != hipSuccess) {
    g.cur = e.cur0;  // nothing of the sweep ran
    g.dir = e.dir0;
    if (rc) return rc;
    return fail(DQMC_E_HIP, std::string("sweep graph: ") + hipGetErrorString(ce != hipSuccess ? ce : ie));
  }
  e.cur1 = g.cur;
  e.dir1 = g.dir;
  g.graphs.push_back(e);
  HIPCHK(hipGraphLaunch(e.exec, g.stream));
  return 0;
}

// ---------------------------------------------------------------------------
// Merged sweeps: every shape group of a batch advanced in lockstep, one launch per kernel for
// all chains. Each group still follows its own propagate state machine (stack.jl:391-499):
// with M a multiple of safe_mult s, step j (1-based, counted from init or from the start of a
// sweep call) recomputes G exactly when j % s == 0 for every chain, whatever its M, so the
// groups always need the same kind of step (a wrap, or a stack rebuild). Where they differ --
// which stack element is rebuilt or loaded, which slice matrices apply, whether a turning
// point resets an element or wraps after the recompute -- each chain reads its own plan row
// (ChainDesc::plan, MK_PLAN references) and per-launch selector / slice rows. A chain that
// has done its sweeps leaves the launches (compacted descriptor arrays per active set).
// ---------------------------------------------------------------------------
enum StepKind : int { SK_WRAP_UP = 0, SK_WRAP_DOWN, SK_UR, SK_DR, SK_BT, SK_TT };

// What one propagate() of a group does (the host twin of Ops::propagate); advances g.cur/dir.
struct StepRec {
  int kind;
  int cur_new;   // current_slice after the step (local updates at cur_new - 1)
  int wslice;    // 0-based slice of the wrap of a wrap step, of UR's drift-check wrap, of DR/TT's final wrap
  int a, b;      // add_slice_sequence: source element a, rebuilt element b
  bool left;     // add_slice_sequence_left (B) or _right (B^dagger)
  int pre, post;  // element loaded before the rebuild (UR, DR), element loaded after it
};

int step_record(Group& g, StepRec& r) {
  const int M = g.M, s = g.s;
  if (g.dir == 1) {
    if (g.cur % s == 0) {
      if (g.cur == 0) return fail(DQMC_E_STATE, "merged sweep: unexpected state (slice 0, up)");
      const int cn = g.cur + 1;
      if (cn <= M) {
        const int idx = (cn - 1) / s;
        r = StepRec{SK_UR, cn, cn - 2, idx - 1, idx, true, idx, idx};
        g.cur = cn;
        return 0;
      }
      r = StepRec{SK_TT, M, M - 1, g.nelem - 2, g.nelem - 1, true, -1, g.nelem - 1};
      g.cur = M;
      g.dir = -1;
      return 0;
    }
    r = StepRec{SK_WRAP_UP, g.cur + 1, g.cur - 1, 0, 0, true, -1, -1};
    g.cur += 1;
    return 0;
  }
  if ((g.cur - 1) % s == 0) {
    const int cn = g.cur - 1;
    if (cn == M) return fail(DQMC_E_STATE, "merged sweep: unexpected state (slice M + 1, down)");
    if (cn > 0) {
      const int idx = cn / s + 1;
      r = StepRec{SK_DR, cn, cn - 1, idx, idx - 1, false, idx - 1, idx - 1};
      g.cur = cn;
      return 0;
    }
    r = StepRec{SK_BT, 1, -1, 1, 0, false, -1, 0};
    g.cur = 1;
    g.dir = 1;
    return 0;
  }
  r = StepRec{SK_WRAP_DOWN, g.cur - 1, g.cur - 2, 0, 0, true, -1, -1};
  g.cur -= 1;
  return 0;
}

// plan-row roles of one recompute step for a chain of group g with record r
void plan_row(const Group& g, const StepRec& r, int* row) {
  for (int k = 0; k < kPlanRoles; ++k) row[k] = -1;
  const int U = MK_USTACK, T = MK_TSTACK, D = VK_DSTACK, NMD = MK_NAMED, NVD = VK_NAMED;
  auto load = [&](int base, int elem, bool to_left) {
    row[base + 0] = plan_enc(NMD, to_left ? M_UL : M_UR);
    row[base + 1] = plan_enc(NMD, to_left ? M_TL : M_TR);
    row[base + 2] = plan_enc(NVD, to_left ? V_DL : V_DR);
    row[base + 3] = plan_enc(U, elem);
    row[base + 4] = plan_enc(T, elem);
    row[base + 5] = plan_enc(D, elem);
  };
  if (r.kind == SK_UR) load(PR_PRE_DU, r.pre, false);
  if (r.kind == SK_DR) load(PR_PRE_DU, r.pre, true);
  row[PR_ADD_SU] = plan_enc(U, r.a);
  row[PR_ADD_SD] = plan_enc(D, r.a);
  row[PR_ADD_ST] = plan_enc(T, r.a);
  row[PR_ADD_DU] = plan_enc(U, r.b);
  row[PR_ADD_DD] = plan_enc(D, r.b);
  row[PR_ADD_DT] = plan_enc(T, r.b);
  const bool post_left = r.kind == SK_UR || r.kind == SK_TT;
  load(PR_P1_DU, r.post, post_left);
  if (r.kind == SK_BT || r.kind == SK_TT) {
    row[PR_ID_U] = plan_enc(U, r.post);
    row[PR_ID_T] = plan_enc(T, r.post);
    row[PR_ID_D] = plan_enc(D, r.post);
    load(PR_P2_DU, r.post, !post_left);
  }
  if (g.all_checks && (r.kind == SK_UR || r.kind == SK_DR)) row[PR_GW] = plan_enc(NMD, M_GW);
}

constexpr int kMergedApplies = 4 + 16;  // apply launches of a step with a selector row: 2 + s + 2, s <= 16

// Shape facts of a merged call and the layout of its plan buffer: per chain [T][kPlanRoles]
// rows, then [T][kMergedApplies][nch] selectors and slices, then [T][nch] local-update slices,
// then per active set the block map of the blocked QR.
constexpr int kQrMapMax = 256;  // cooperative blocks of one QR launch: one per CU

// Workgroups per chain for the blocked QR of one launch over chains of sizes n[k]: every chain
// starts at 1, then the chain with the longest estimated factorisation takes one more until
// kQrMapMax are given out (at most n / 8 and kMaxWG each). Estimate, fitted to single-chain
// timings on MI355X at n = 256 ... 576: t(n, W) = c(n) + d(n) / (W - 1) for W >= 2 (the panel
// chain of workgroup 0, and the trailing update shared by the other W - 1), c ~ n^1.24,
// d ~ n^3.83; with W = 1 the one workgroup does both, 2.26 (c + d). The map lists
// chain | w << 16 | W << 24 per block.
int qr_block_map(const std::vector<int>& n, std::vector<int>& map) {
  const int na = (int)n.size();
  if (na > kQrMapMax) return -1;
  auto est = [](int nn, int W) {
    const double x = nn / 400.0, c = 0.7 * std::pow(x, 1.24), d = 4.2 * std::pow(x, 3.83);
    return W >= 2 ? c + d / (W - 1) : 2.26 * (c + d);
  };
  std::vector<int> W(na, 1);
  for (int sum = na; sum < kQrMapMax; ++sum) {
    int best = -1;
    double bt = 0.0;
    for (int k = 0; k < na; ++k)
      if (W[k] < std::max(1, std::min(n[k] / 8, kMaxWG))) {
        const double t = est(n[k], W[k]);
        if (t > bt) {
          bt = t;
          best = k;
        }
      }
    if (best < 0) break;
    ++W[best];
  }
  map.clear();
  for (int k = 0; k < na; ++k)
    for (int w = 0; w < W[k]; ++w) map.push_back(k | (w << 16) | (W[k] << 24));
  return (int)map.size();
}

// Shape facts of a merged call and the layout of its plan buffer: per chain [T][kPlanRoles]
// rows, then [T][kMergedApplies][nch] selectors and slices, then [T][nch] local-update slices.
struct MergedShape {
  int s = 0, Mmax = 0, nch = 0;
  std::vector<int> Ms;                // distinct M, ascending
  std::vector<std::vector<int>> sets;  // set k: the chains with M >= Ms[k] (active while t < nsweeps Ms[k])
  std::vector<std::vector<int>> qr_maps;  // set k: its QR block map (at qr_map_off(T) + k * kQrMapMax)
  std::vector<std::vector<int>> local_maps;  // set k: local-update block map (after the QR maps)
  size_t qr_map_off(int T) const {
    return (size_t)nch * T * kPlanRoles + 2 * (size_t)T * kMergedApplies * nch + (size_t)T * nch;
  }
  size_t local_map_off(int T) const { return qr_map_off(T) + sets.size() * (size_t)kQrMapMax; }
  size_t ints(int T) const { return local_map_off(T) + sets.size() * (size_t)kQrMapMax; }
};

MergedShape merged_shape(dqmc_batch* b) {
  MergedShape ms;
  ms.s = b->groups[0].s;
  ms.nch = (int)b->chains.size();
  for (auto& g : b->groups) {
    ms.Mmax = std::max(ms.Mmax, g.M);
    ms.Ms.push_back(g.M);
  }
  std::sort(ms.Ms.begin(), ms.Ms.end());
  ms.Ms.erase(std::unique(ms.Ms.begin(), ms.Ms.end()), ms.Ms.end());
  for (int Mk : ms.Ms) {
    std::vector<int> v;
    for (int ci = 0; ci < ms.nch; ++ci)
      if (b->chains[ci].M >= Mk) v.push_back(ci);
    ms.sets.push_back(v);
    std::vector<int> ns, map;
    for (int ci : v) ns.push_back(b->chains[ci].n);
    if (qr_block_map(ns, map) < 0) map.clear();  // more chains than blocks: the uniform grid
    ms.qr_maps.push_back(map);
    if (local_block_map(ns, map) < 0) map.clear();
    ms.local_maps.push_back(map);
  }
  return ms;
}

// The host side of steps t0 .. t0+T-1 of a call of nsweeps sweeps (advances the groups' states):
// plan ints (layout of MergedShape::ints(T)), the step kind (0 wrap, 1 recompute) and active set.
int merged_plan(dqmc_batch* b, const MergedShape& ms, int nsweeps, long long t0, int T, std::vector<int>& h,
                std::vector<int>& kinds, std::vector<int>& set_of) {
  const int ng = (int)b->groups.size(), nch = ms.nch, s = ms.s;
  h.assign(ms.ints(T), 0);
  int* hrows = h.data();
  int* hsel = hrows + (size_t)nch * T * kPlanRoles;
  int* hsli = hsel + (size_t)T * kMergedApplies * nch;
  int* hloc = hsli + (size_t)T * kMergedApplies * nch;
  std::fill(hrows, hsel, -1);
  std::fill(hsel, hsli, -1);
  for (size_t k = 0; k < ms.sets.size(); ++k) {
    std::copy(ms.qr_maps[k].begin(), ms.qr_maps[k].end(), h.begin() + ms.qr_map_off(T) + k * kQrMapMax);
    std::copy(ms.local_maps[k].begin(), ms.local_maps[k].end(), h.begin() + ms.local_map_off(T) + k * kQrMapMax);
  }
  kinds.assign(T, 0);
  set_of.assign(T, 0);
  std::vector<StepRec> rec(ng);
  for (int t = 0; t < T; ++t) {
    const long long ts = t0 + t;
    int sk = 0;
    while (sk < (int)ms.Ms.size() && ts >= (long long)nsweeps * ms.Ms[sk]) ++sk;
    set_of[t] = sk;
    const std::vector<int>& act = ms.sets[sk];
    int kind_cls = -1;
    for (int gi = 0; gi < ng; ++gi) {
      Group& g = b->groups[gi];
      if (ts >= (long long)nsweeps * g.M) continue;
      RC(step_record(g, rec[gi]));
      const int cls = rec[gi].kind <= SK_WRAP_DOWN ? 0 : 1;
      if (kind_cls < 0) kind_cls = cls;
      if (cls != kind_cls) return fail(DQMC_E_STATE, "merged sweep: groups out of step (safe_mult / slices)");
    }
    kinds[t] = kind_cls;
    for (int k = 0; k < (int)act.size(); ++k) {
      const int ci = act[k];
      const Group& g = b->groups[b->chain_group[ci]];
      const StepRec& r = rec[b->chain_group[ci]];
      int* row = hrows + ((size_t)ci * T + t) * kPlanRoles;
      int* sel = hsel + (size_t)t * kMergedApplies * nch;
      int* sli = hsli + (size_t)t * kMergedApplies * nch;
      hloc[(size_t)t * nch + k] = r.cur_new - 1;
      auto set_apply = [&](int a, int which, int slice) {
        sel[(size_t)a * nch + k] = which;
        sli[(size_t)a * nch + k] = slice;
      };
      if (kind_cls == 0) {
        const int w = r.kind == SK_WRAP_UP ? 0 : 1;
        set_apply(0, w, r.wslice);
        set_apply(1, w, r.wslice);
        continue;
      }
      plan_row(g, r, row);
      if (r.kind == SK_UR && g.all_checks) {  // drift check: wrap(GW, cur_new - 1, +1)
        set_apply(0, 0, r.wslice);
        set_apply(1, 0, r.wslice);
      }
      for (int j = 0; j < s; ++j)  // add_slice_sequence_left: B_l ascending; _right: B_l^dagger descending
        set_apply(2 + j, r.left ? 0 : 1, r.left ? r.a * s + j : r.a * s - 1 - j);
      if (r.kind == SK_DR || r.kind == SK_TT) {  // wrap(G, cur_new + 1, -1)
        set_apply(2 + s, 0, r.wslice);
        set_apply(3 + s, 0, r.wslice);
      }
    }
  }
  return 0;
}

// Descriptor arrays of the active sets, each chain's pointing to its plan rows at d_plan.
void merged_descs(dqmc_batch* b, const MergedShape& ms, int T, const int* d_plan, std::vector<ChainDesc>& h,
                  std::vector<size_t>& set_off) {
  h.clear();
  set_off.assign(ms.sets.size(), 0);
  for (size_t k = 0; k < ms.sets.size(); ++k) {
    set_off[k] = h.size();
    for (int ci : ms.sets[k]) {
      ChainDesc d = b->chains[ci].d;
      d.plan = d_plan + (size_t)ci * T * kPlanRoles;
      h.push_back(d);
    }
  }
}

// Enqueue the launches of T planned steps on b->mstream (no host synchronisation).
int merged_enqueue(dqmc_batch* b, const MergedShape& ms, int T, const int* d_plan, const ChainDesc* d_desc,
                   const std::vector<size_t>& set_off, const std::vector<int>& kinds, const std::vector<int>& set_of,
                   double box, int dolog) {
  const Group& g0 = b->groups[0];
  const int nch = ms.nch, s = ms.s;
  const int* d_sel = d_plan + (size_t)nch * T * kPlanRoles;
  const int* d_sli = d_sel + (size_t)T * kMergedApplies * nch;
  const int* d_loc = d_sli + (size_t)T * kMergedApplies * nch;
  Group mg;  // a pseudo group for the Ops helpers: the merged stream, the current active set
  mg.L = 0;
  mg.M = ms.Mmax;
  mg.s = s;
  mg.nelem = 0;
  mg.chkr = g0.chkr;
  mg.ngroups = g0.ngroups;
  mg.all_checks = g0.all_checks;
  mg.stream = b->mstream;
  mg.greens_route = g0.greens_route;
  mg.udt_route = g0.udt_route;
  mg.qr_nb = g0.qr_nb ? g0.qr_nb : 8;
  mg.wg_inv = g0.wg_inv;
  mg.local_variant = g0.local_variant;
  mg.prof_mask = g0.prof_mask;
  for (int t = 0; t < T; ++t) {
    const std::vector<int>& act = ms.sets[set_of[t]];
    const int na = (int)act.size();
    int an = 0, aN = 0, amin = 1 << 30;
    for (int ci : act) {
      an = std::max(an, b->chains[ci].n);
      aN = std::max(aN, b->chains[ci].N);
      amin = std::min(amin, b->chains[ci].n);
    }
    mg.d_desc = const_cast<ChainDesc*>(d_desc) + set_off[set_of[t]];
    const std::vector<int>& qmap = ms.qr_maps[set_of[t]];
    mg.qr_map = qmap.empty() ? nullptr : d_plan + ms.qr_map_off(T) + (size_t)set_of[t] * kQrMapMax;
    mg.qr_map_blocks = (int)qmap.size();
    const std::vector<int>& lmap = ms.local_maps[set_of[t]];
    mg.local_map = lmap.empty() ? nullptr : d_plan + ms.local_map_off(T) + (size_t)set_of[t] * kQrMapMax;
    mg.local_map_blocks = (int)lmap.size();
    mg.chains.assign(na, 0);
    mg.n = an;
    mg.N = aN;
    // one workgroup per CU over all active chains
    const int per = std::max(1, 256 / na);
    mg.wg_qr = std::max(1, std::min({per, amin / mg.qr_nb, kMaxWG}));
    mg.wg_local = std::max(1, std::min({per, amin / 4, kMaxWG}));
    if (local_window_sites(an, mg.wg_local) < 2) mg.wg_local = std::max(1, std::min({per, amin / 8, kMaxWG}));
    Ops o(mg);
    const ChainDesc* dd = mg.d_desc;
    const int* sel = d_sel + (size_t)t * kMergedApplies * nch;
    const int* sli = d_sli + (size_t)t * kMergedApplies * nch;
    auto apply = [&](MatRef X, BOp op0, BOp op1, int a) -> int {
      ProfScope ps(mg, DQMC_K_APPLY_B);
      LCHK(launch_apply_B_chkr_sel(mg.stream, dd, na, an, aN, X, op0, op1, sel + (size_t)a * nch, sli + (size_t)a * nch,
                                   mg.ngroups));
      return 0;
    };
    auto P = [&](int role) { return MatRef{MK_PLAN, t * kPlanRoles + role}; };
    auto PV = [&](int role) { return VecRef{VK_PLAN, t * kPlanRoles + role}; };
    if (kinds[t] == 0) {
      RC(apply(NM(M_G), B_LEFT, B_INV_LEFT, 0));
      RC(apply(NM(M_G), B_INV_RIGHT, B_RIGHT, 1));
    } else {
      RC(o.copy(P(PR_PRE_DU), P(PR_PRE_SU)));
      RC(o.copy(P(PR_PRE_DT), P(PR_PRE_ST)));
      RC(o.copyv(PV(PR_PRE_DD), PV(PR_PRE_SD)));
      if (mg.all_checks) {
        RC(o.copy(P(PR_GW), NM(M_G)));
        RC(apply(NM(M_GW), B_LEFT, B_LEFT, 0));
        RC(apply(NM(M_GW), B_INV_RIGHT, B_INV_RIGHT, 1));
      }
      RC(o.copy(NM(M_CURU), P(PR_ADD_SU)));
      for (int j = 0; j < s; ++j) RC(apply(NM(M_CURU), B_LEFT, B_DAG_LEFT, 2 + j));
      LCHK(launch_scale_cols(mg.stream, dd, na, an, NM(M_CURU), PV(PR_ADD_SD)));
      RC(o.udt(NM(M_CURU), PV(PR_ADD_DD), P(PR_ADD_DU), NM(M_QRT), true));
      RC(o.gemm(P(PR_ADD_DT), NM(M_QRT), OP_N, P(PR_ADD_ST), OP_N));
      RC(o.copy(P(PR_P1_DU), P(PR_P1_SU)));
      RC(o.copy(P(PR_P1_DT), P(PR_P1_ST)));
      RC(o.copyv(PV(PR_P1_DD), PV(PR_P1_SD)));
      LCHK(launch_identity(mg.stream, dd, na, an, P(PR_ID_U), NOVEC));
      LCHK(launch_identity(mg.stream, dd, na, an, P(PR_ID_T), NOVEC));
      LCHK(launch_ones_vec(mg.stream, dd, na, an, PV(PR_ID_D)));
      RC(o.copy(P(PR_P2_DU), P(PR_P2_SU)));
      RC(o.copy(P(PR_P2_DT), P(PR_P2_ST)));
      RC(o.copyv(PV(PR_P2_DD), PV(PR_P2_SD)));
      RC(o.calculate_greens(true));
      if (mg.all_checks) LCHK(launch_maxdiff(mg.stream, dd, na, an, P(PR_GW), NM(M_G)));
      RC(apply(NM(M_G), B_INV_LEFT, B_INV_LEFT, 2 + s));
      RC(apply(NM(M_G), B_RIGHT, B_RIGHT, 3 + s));
    }
    {
      ProfScope ps(mg, DQMC_K_LOCAL);
      int variant = mg.local_variant;
      if (variant == 0)
        variant = local_window_sites(an, mg.wg_local) >= 2 ? 4 : local_tile_fits(an, mg.wg_local) ? 2
                  : local_delayed_fits(an) ? 3 : 1;
      // without LDS-resident columns the chains of a launch get workgroups by size
      bool tile = true;
      if (variant == 4) local_window_sites(an, mg.wg_local, &tile);
      const bool by_size = variant == 4 && !tile && mg.local_map;
      LCHK(launch_local_updates(mg.stream, dd, na, an, 0, d_loc + (size_t)t * nch, box, mg.wg_local, dolog, variant,
                                by_size ? mg.local_map : nullptr, by_size ? mg.local_map_blocks : 0));
    }
  }
  // timing records of the pseudo group belong to the first group (dqmc_batch_profile_read)
  if (!mg.prof.empty()) {
    Group& gp0 = b->groups[0];
    gp0.prof.insert(gp0.prof.end(), mg.prof.begin(), mg.prof.end());
    mg.prof.clear();
  }
  return 0;
}

// the merged stream waits for the groups' queued work, and later group work for it
int merged_enter(dqmc_batch* b) {
  for (auto& g : b->groups) {
    hipEvent_t e = g.take_event();
    HIPCHK(hipEventRecord(e, g.stream));
    HIPCHK(hipStreamWaitEvent(b->mstream, e, 0));
    g.event_pool.push_back(e);
  }
  return 0;
}
int merged_leave(dqmc_batch* b) {
  hipEvent_t e = nullptr;
  HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  HIPCHK(hipEventRecord(e, b->mstream));
  for (auto& g : b->groups) HIPCHK(hipStreamWaitEvent(g.stream, e, 0));
  (void)hipEventDestroy(e);
  return 0;
}

// Enqueue nsweeps sweeps of every chain of the batch (all groups compatible) on b->mstream,
// planned in chunks of Mmax steps through one reusable plan buffer.
int merged_sweeps(dqmc_batch* b, int nsweeps, double box, int dolog) {
  const MergedShape ms = merged_shape(b);
  if (ms.s > 16) return fail(DQMC_E_ARG, "merged sweeps need safe_mult <= 16");
  const long long total = (long long)nsweeps * ms.Mmax;
  if (total == 0) return 0;
  const int TC = (int)std::min<long long>(total, ms.Mmax);
  const size_t need = ms.ints(TC);
  HIPCHK(hipStreamSynchronize(b->mstream));  // the previous call's buffers and host rows are free
  if (need > b->mplan_cap) {
    if (b->d_mplan) HIPCHK(hipFree(b->d_mplan));
    b->d_mplan = nullptr;
    HIPCHK(hipMalloc(&b->d_mplan, need * sizeof(int)));
    b->mplan_cap = need;
  }
  size_t ndesc = 0;
  for (auto& v : ms.sets) ndesc += v.size();
  if (ndesc > b->mdesc_cap) {
    if (b->d_mdesc) HIPCHK(hipFree(b->d_mdesc));
    b->d_mdesc = nullptr;
    HIPCHK(hipMalloc(&b->d_mdesc, ndesc * sizeof(ChainDesc)));
    b->mdesc_cap = ndesc;
  }
  std::vector<size_t> set_off;
  merged_descs(b, ms, TC, b->d_mplan, b->h_mdesc, set_off);
  RC(merged_enter(b));
  HIPCHK(hipMemcpyAsync(b->d_mdesc, b->h_mdesc.data(), b->h_mdesc.size() * sizeof(ChainDesc), hipMemcpyHostToDevice,
                        b->mstream));
  std::vector<int> kinds, set_of;
  for (long long t0 = 0; t0 < total; t0 += TC) {
    const int T = (int)std::min<long long>(TC, total - t0);
    // the host rows are copied asynchronously: the previous chunk's copy must be done
    HIPCHK(hipStreamSynchronize(b->mstream));
    RC(merged_plan(b, ms, nsweeps, t0, TC, b->h_rows, kinds, set_of));
    HIPCHK(hipMemcpyAsync(b->d_mplan, b->h_rows.data(), need * sizeof(int), hipMemcpyHostToDevice, b->mstream));
    RC(merged_enqueue(b, ms, T, b->d_mplan, b->d_mdesc, set_off, kinds, set_of, box, dolog));
  }
  return merged_leave(b);
}

// One merged sweep (every chain one sweep) replayed from a HIP graph. The launches are fixed
// by the groups' (current_slice, direction) at the start, the box and the logging flag; each
// captured sweep owns its plan and descriptor buffers.
int merged_sweep_graph(dqmc_batch* b, double box, int dolog) {
  std::vector<int> key;
  for (auto& g : b->groups) {
    key.push_back(g.cur);
    key.push_back(g.dir);
  }
  // the graphs read copies of the descriptors: a changed descriptor (a new uniform buffer or
  // count, measurement or log buffers) is copied into every graph's array before a replay
  std::vector<ChainDesc> descs;
  for (auto& cm : b->chains) descs.push_back(cm.d);
  for (auto& e : b->mgraphs)
    if (std::memcmp(e.descs.data(), descs.data(), descs.size() * sizeof(ChainDesc)) != 0) {
      HIPCHK(hipStreamSynchronize(b->mstream));
      const MergedShape ms = merged_shape(b);
      std::vector<ChainDesc> hd;
      std::vector<size_t> set_off;
      merged_descs(b, ms, ms.Mmax, e.d_plan, hd, set_off);
      HIPCHK(hipMemcpy(e.d_desc, hd.data(), hd.size() * sizeof(ChainDesc), hipMemcpyHostToDevice));
      e.descs = descs;
    }
  for (auto& e : b->mgraphs)
    if (e.key == key && e.box == box && e.dolog == dolog) {
      RC(merged_enter(b));
      HIPCHK(hipGraphLaunch(e.exec, b->mstream));
      for (size_t gi = 0; gi < b->groups.size(); ++gi) {
        b->groups[gi].cur = e.end[2 * gi];
        b->groups[gi].dir = e.end[2 * gi + 1];
      }
      return merged_leave(b);
    }
  const MergedShape ms = merged_shape(b);
  if (ms.s > 16) return fail(DQMC_E_ARG, "merged sweeps need safe_mult <= 16");
  const int T = ms.Mmax;
  std::vector<int> h, kinds, set_of;
  std::vector<ChainDesc> hd;
  std::vector<size_t> set_off;
  const std::vector<int> start = key;
  const int rc0 = merged_plan(b, ms, 1, 0, T, h, kinds, set_of);
  auto restore = [&]() {
    for (size_t gi = 0; gi < b->groups.size(); ++gi) {
      b->groups[gi].cur = start[2 * gi];
      b->groups[gi].dir = start[2 * gi + 1];
    }
  };
  if (rc0) {
    restore();
    return rc0;
  }
  dqmc_batch::MergedGraph e;
  e.key = key;
  e.descs = descs;
  e.box = box;
  e.dolog = dolog;
  for (auto& g : b->groups) {
    e.end.push_back(g.cur);
    e.end.push_back(g.dir);
  }
  size_t ndesc = 0;
  for (auto& v : ms.sets) ndesc += v.size();
  hipError_t me = hipMalloc(&e.d_plan, h.size() * sizeof(int));
  if (me == hipSuccess) me = hipMalloc(&e.d_desc, ndesc * sizeof(ChainDesc));
  if (me == hipSuccess) me = hipMemcpy(e.d_plan, h.data(), h.size() * sizeof(int), hipMemcpyHostToDevice);
  if (me == hipSuccess) {
    merged_descs(b, ms, T, e.d_plan, hd, set_off);
    me = hipMemcpy(e.d_desc, hd.data(), hd.size() * sizeof(ChainDesc), hipMemcpyHostToDevice);
  }
  if (me != hipSuccess) {
    if (e.d_plan) (void)hipFree(e.d_plan);
    if (e.d_desc) (void)hipFree(e.d_desc);
    restore();
    return fail(DQMC_E_HIP, std::string("merged sweep graph buffers: ") + hipGetErrorString(me));
  }
  RC(merged_enter(b));
  HIPCHK(hipStreamSynchronize(b->mstream));
  HIPCHK(hipStreamBeginCapture(b->mstream, hipStreamCaptureModeRelaxed));
  const int rc = merged_enqueue(b, ms, T, e.d_plan, e.d_desc, set_off, kinds, set_of, box, dolog);
  hipGraph_t graph = nullptr;
  const hipError_t ce = hipStreamEndCapture(b->mstream, &graph);
  hipError_t ie = hipSuccess;
  if (!rc && ce == hipSuccess) ie = hipGraphInstantiate(&e.exec, graph, nullptr, nullptr, 0);
  if (graph) (void)hipGraphDestroy(graph);
  if (rc || ce != hipSuccess || ie != hipSuccess) {
    (void)hipFree(e.d_plan);
    (void)hipFree(e.d_desc);
    restore();
    if (rc) return rc;
    return fail(DQMC_E_HIP, std::string("merged sweep graph: ") + hipGetErrorString(ce != hipSuccess ? ce : ie));
  }
  b->mgraphs.push_back(e);
  HIPCHK(hipGraphLaunch(e.exec, b->mstream));
  return merged_leave(b);
}

// device addresses of a chain's matrices and vectors, as matptr/vecptr compute them on the device
cplx* host_matptr(const ChainMem& cm, MatRef r) {
  const size_t nn = (size_t)cm.n * cm.n;
  switch (r.kind) {
    case MK_NAMED: return cm.d.mat[r.idx];
    case MK_USTACK: return cm.d.ustack + nn * r.idx;
    case MK_TSTACK: return cm.d.tstack + nn * r.idx;
    default: return cm.d.xmat + nn * r.idx;
  }
}
double* host_vecptr(const ChainMem& cm, VecRef r) {
  switch (r.kind) {
    case VK_NAMED: return cm.d.vec[r.idx];
    case VK_DSTACK: return cm.d.dstack + (size_t)cm.n * r.idx;
    default: return cm.d.xvec + (size_t)cm.n * r.idx;
  }
}

// host <-> device matrix helpers for the unit surfaces of a single-chain context
int put_mat(dqmc_batch* b, MatRef r, const double* host) {
  ChainMem& cm = b->chains[0];
  const size_t nn = (size_t)cm.n * cm.n;
  HIPCHK(hipMemcpyAsync(host_matptr(cm, r), host, nn * sizeof(cplx), hipMemcpyHostToDevice, b->groups[0].stream));
  return 0;
}
int get_mat(dqmc_batch* b, MatRef r, double* host) {
  ChainMem& cm = b->chains[0];
  const size_t nn = (size_t)cm.n * cm.n;
  HIPCHK(hipMemcpyAsync(host, host_matptr(cm, r), nn * sizeof(cplx), hipMemcpyDeviceToHost, b->groups[0].stream));
  HIPCHK(hipStreamSynchronize(b->groups[0].stream));
  return 0;
}
int put_vec(dqmc_batch* b, VecRef r, const double* host) {
  ChainMem& cm = b->chains[0];
  HIPCHK(hipMemcpyAsync(host_vecptr(cm, r), host, cm.n * sizeof(double), hipMemcpyHostToDevice,
                        b->groups[0].stream));
  return 0;
}
int get_vec(dqmc_batch* b, VecRef r, double* host) {
  ChainMem& cm = b->chains[0];
  HIPCHK(hipMemcpyAsync(host, host_vecptr(cm, r), cm.n * sizeof(double), hipMemcpyDeviceToHost,
                        b->groups[0].stream));
  HIPCHK(hipStreamSynchronize(b->groups[0].stream));
  return 0;
}

double read_scalar(dqmc_batch* b, int chain, int slot) {
  double v = 0;
  (void)hipMemcpy(&v, b->chains[chain].d.scal + slot, sizeof(double), hipMemcpyDeviceToHost);
  return v;
}

}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================
#define CTX_CHECK(ctx) \
  if (!(ctx) || !(ctx)->b) return fail(DQMC_E_ARG, "null context")

extern "C" {

const char* dqmc_last_error(void) { return g_err.c_str(); }

int dqmc_ctx_create(const dqmc_params* p, const dqmc_operators* ops, int device, dqmc_ctx** out) {
  if (!out) return fail(DQMC_E_ARG, "null out");
  dqmc_batch* b = nullptr;
  RC(batch_create(p, ops, 1, device, &b));
  *out = new dqmc_ctx{b};
  return 0;
}

int dqmc_ctx_destroy(dqmc_ctx* ctx) {
  if (!ctx) return 0;
  batch_destroy(ctx->b);
  delete ctx;
  return 0;
}

int dqmc_batch_set_groups(dqmc_batch* b, int wg_qr, int wg_inv, int wg_local) {
  if (!b) return fail(DQMC_E_ARG, "null batch");
  if (wg_qr > kMaxWG || wg_inv > kMaxWG || wg_local > kMaxWG) return fail(DQMC_E_ARG, "too many workgroups");
  RC(batch_sync(b));
  b->drop_mgraphs();
  for (auto& g : b->groups) {
    g.drop_graphs();
    if (wg_qr > 0) g.wg_qr = wg_qr;
    if (wg_inv > 0) g.wg_inv = std::max(wg_inv, (g.n + 511) / 512);
    if (wg_local > 0) g.wg_local = wg_local;
  }
  return 0;
}

int dqmc_ctx_set_groups(dqmc_ctx* ctx, int wg_qr, int wg_inv, int wg_local) {
  CTX_CHECK(ctx);
  return dqmc_batch_set_groups(ctx->b, wg_qr, wg_inv, wg_local);
}

int dqmc_set_greens_route(dqmc_ctx* ctx, int route) {
  CTX_CHECK(ctx);
  if (route != 0 && route != 1) return fail(DQMC_E_ARG, "route must be 0 or 1");
  RC(batch_sync(ctx->b));
  ctx->b->drop_mgraphs();
  ctx->b->groups[0].drop_graphs();
  ctx->b->groups[0].greens_route = route;
  return 0;
}

int dqmc_set_udt_route(dqmc_ctx* ctx, int route) {
  CTX_CHECK(ctx);
  if (route != 0 && route != 1) return fail(DQMC_E_ARG, "udt route must be 0 or 1");
  RC(batch_sync(ctx->b));
  ctx->b->drop_mgraphs();
  for (auto& g : ctx->b->groups) {
    g.drop_graphs();
    g.udt_route = route;
  }
  return 0;
}

int dqmc_batch_set_local_variant(dqmc_batch* b, int variant) {
  if (!b) return fail(DQMC_E_ARG, "null batch");
  if (variant < 0 || variant > 4) return fail(DQMC_E_ARG, "local variant must be 0, 1, 2, 3 or 4");
  RC(batch_sync(b));
  b->drop_mgraphs();
  for (auto& g : b->groups) {
    g.drop_graphs();
    g.local_variant = variant;
  }
  return 0;
}

int dqmc_batch_set_qr_panel(dqmc_batch* b, int nb) {
  if (!b) return fail(DQMC_E_ARG, "null batch");
  if (nb != 0 && nb != 8 && nb != 16) return fail(DQMC_E_ARG, "QR panel width must be 0, 8 or 16");
  for (auto& g : b->groups)
    if (nb == 16 && g.n > 512) return fail(DQMC_E_ARG, "16-column QR panels need n <= 512");
  RC(batch_sync(b));
  b->drop_mgraphs();
  for (auto& g : b->groups) {
    g.drop_graphs();
    g.qr_nb = nb ? nb : 8;  // the default of default_groups
  }
  return 0;
}

int dqmc_set_qr_panel(dqmc_ctx* ctx, int nb) {
  CTX_CHECK(ctx);
  return dqmc_batch_set_qr_panel(ctx->b, nb);
}

int dqmc_set_local_variant(dqmc_ctx* ctx, int variant) {
  CTX_CHECK(ctx);
  return dqmc_batch_set_local_variant(ctx->b, variant);
}

int dqmc_set_field(dqmc_ctx* ctx, const double* phi) {
  CTX_CHECK(ctx);
  return dqmc_batch_set_field(ctx->b, 0, phi);
}
int dqmc_get_field(dqmc_ctx* ctx, double* phi) {
  CTX_CHECK(ctx);
  return dqmc_batch_get_field(ctx->b, 0, phi);
}
int dqmc_set_uniforms(dqmc_ctx* ctx, const double* u, long long count) {
  CTX_CHECK(ctx);
  return set_uniforms(ctx->b, 0, u, count);
}

int dqmc_init(dqmc_ctx* ctx) {
  CTX_CHECK(ctx);
  return dqmc_batch_init(ctx->b);
}

int dqmc_build_stack(dqmc_ctx* ctx) {
  CTX_CHECK(ctx);
  Ops o(ctx->b->groups[0]);
  return o.build_stack();
}

int dqmc_propagate(dqmc_ctx* ctx, int* current_slice, int* direction, double* drift) {
  CTX_CHECK(ctx);
  Group& g = ctx->b->groups[0];
  Ops o(g);
  RC(o.propagate());
  if (current_slice) *current_slice = g.cur;
  if (direction) *direction = g.dir;
  if (drift) {
    RC(batch_sync(ctx->b));
    *drift = g.recomputed ? read_scalar(ctx->b, 0, S_DRIFT_LAST) : -1.0;
  }
  return 0;
}

int dqmc_local_updates(dqmc_ctx* ctx, double box, double* acc_rate) {
  CTX_CHECK(ctx);
  Group& g = ctx->b->groups[0];
  Ops o(g);
  const int dolog = ctx->b->chains[0].log_cap > 0;
  RC(o.local_updates(box, dolog));
  if (acc_rate) {
    RC(batch_sync(ctx->b));
    *acc_rate = read_scalar(ctx->b, 0, S_ACC_SLICE) / ctx->b->chains[0].N;
  }
  return 0;
}

int dqmc_sweep(dqmc_ctx* ctx, int nsweeps, double box) {
  CTX_CHECK(ctx);
  return dqmc_batch_sweep(ctx->b, nsweeps, &box);
}

int dqmc_synchronize(dqmc_ctx* ctx) {
  CTX_CHECK(ctx);
  return batch_sync(ctx->b);
}

int dqmc_get_greens(dqmc_ctx* ctx, double* G) {
  CTX_CHECK(ctx);
  return dqmc_batch_get_greens(ctx->b, 0, G);
}

int dqmc_set_greens(dqmc_ctx* ctx, const double* G) {
  CTX_CHECK(ctx);
  if (!G) return fail(DQMC_E_ARG, "null G");
  RC(put_mat(ctx->b, NM(M_G), G));
  return batch_sync(ctx->b);
}

int dqmc_get_stack(dqmc_ctx* ctx, int k, double* U, double* D, double* T) {
  CTX_CHECK(ctx);
  ChainMem& cm = ctx->b->chains[0];
  if (k < 0 || k >= cm.nelem) return fail(DQMC_E_ARG, "stack index out of range");
  RC(batch_sync(ctx->b));
  if (U) RC(get_mat(ctx->b, US(k), U));
  if (T) RC(get_mat(ctx->b, TS(k), T));
  if (D) RC(get_vec(ctx->b, DS(k), D));
  return 0;
}

int dqmc_get_stats(dqmc_ctx* ctx, dqmc_stats* st) {
  CTX_CHECK(ctx);
  return get_stats(ctx->b, 0, st);
}

int dqmc_reset_stats(dqmc_ctx* ctx) {
  CTX_CHECK(ctx);
  Group& g = ctx->b->groups[0];
  LCHK(launch_reset_stats(g.stream, g.d_desc, g.nch()));
  return batch_sync(ctx->b);
}

int dqmc_set_proposal_log(dqmc_ctx* ctx, long long capacity) {
  CTX_CHECK(ctx);
  dqmc_batch* b = ctx->b;
  RC(batch_sync(b));
  ChainMem& cm = b->chains[0];
  if (cm.log_buf) HIPCHK(hipFree(cm.log_buf));
  cm.log_buf = nullptr;
  cm.log_cap = 0;
  if (capacity > 0) {
    HIPCHK(hipMalloc(&cm.log_buf, sizeof(double) * 4 * capacity));
    cm.log_cap = capacity;
  }
  cm.d.prop_log = cm.log_buf;
  cm.d.log_cap = cm.log_cap;
  long long zero = 0;
  HIPCHK(hipMemcpy(cm.d.log_count, &zero, sizeof(zero), hipMemcpyHostToDevice));
  return upload_desc(b, b->groups[0]);
}

int dqmc_get_proposal_log(dqmc_ctx* ctx, double* out, long long* count) {
  CTX_CHECK(ctx);
  dqmc_batch* b = ctx->b;
  RC(batch_sync(b));
  ChainMem& cm = b->chains[0];
  long long c = 0;
  HIPCHK(hipMemcpy(&c, cm.d.log_count, sizeof(c), hipMemcpyDeviceToHost));
  if (count) *count = c;
  const long long m = std::min(c, cm.log_cap);
  if (out && m > 0) HIPCHK(hipMemcpy(out, cm.log_buf, sizeof(double) * 4 * m, hipMemcpyDeviceToHost));
  return 0;
}

int dqmc_apply_B(dqmc_ctx* ctx, int slice, int which, double* X) {
  CTX_CHECK(ctx);
  ChainMem& cm = ctx->b->chains[0];
  if (!X || slice < 0 || slice >= cm.M || which < 0 || which > 4) return fail(DQMC_E_ARG, "bad apply_B args");
  Ops o(ctx->b->groups[0]);
  RC(put_mat(ctx->b, NM(M_GRAD), X));
  RC(o.applyB(NM(M_GRAD), (BOp)which, slice));
  RC(get_mat(ctx->b, NM(M_GRAD), X));
  return check_device_error(ctx->b);
}

int dqmc_wrap_greens(dqmc_ctx* ctx, double* G, int slice, int direction) {
  CTX_CHECK(ctx);
  ChainMem& cm = ctx->b->chains[0];
  if (!G || (direction != 1 && direction != -1)) return fail(DQMC_E_ARG, "bad wrap args");
  const int sl0 = direction == 1 ? slice - 1 : slice - 2;
  if (sl0 < 0 || sl0 >= cm.M) return fail(DQMC_E_ARG, "slice out of range");
  Ops o(ctx->b->groups[0]);
  RC(put_mat(ctx->b, NM(M_GRAD), G));
  RC(o.wrap(NM(M_GRAD), slice, direction));
  return get_mat(ctx->b, NM(M_GRAD), G);
}

int dqmc_udt(dqmc_ctx* ctx, const double* A, double* U, double* D, double* T) {
  CTX_CHECK(ctx);
  if (!A) return fail(DQMC_E_ARG, "null A");
  Ops o(ctx->b->groups[0]);
  RC(put_mat(ctx->b, NM(M_QRA), A));
  RC(o.udt(NM(M_QRA), NV(V_QRD), NM(M_QRQ), NM(M_QRT)));
  RC(batch_sync(ctx->b));
  if (U) RC(get_mat(ctx->b, NM(M_QRQ), U));
  if (T) RC(get_mat(ctx->b, NM(M_QRT), T));
  if (D) RC(get_vec(ctx->b, NV(V_QRD), D));
  return 0;
}

int dqmc_udt_prepivoted(dqmc_ctx* ctx, const double* A, double* U, double* D, double* T) {
  CTX_CHECK(ctx);
  if (!A) return fail(DQMC_E_ARG, "null A");
  Group& g = ctx->b->groups[0];
  const int saved = g.udt_route;
  g.udt_route = 0;
  Ops o(g);
  RC(put_mat(ctx->b, NM(M_QRA), A));
  const int rc = o.udt(NM(M_QRA), NV(V_QRD), NM(M_QRQ), NM(M_QRT), true);
  g.udt_route = saved;
  RC(rc);
  RC(batch_sync(ctx->b));
  if (U) RC(get_mat(ctx->b, NM(M_QRQ), U));
  if (T) RC(get_mat(ctx->b, NM(M_QRT), T));
  if (D) RC(get_vec(ctx->b, NV(V_QRD), D));
  return 0;
}

int dqmc_calculate_greens(dqmc_ctx* ctx, const double* Ul, const double* Dl, const double* Tl, const double* Ur,
                          const double* Dr, const double* Tr, double* G, double* logdet) {
  CTX_CHECK(ctx);
  if (!Ul || !Dl || !Tl || !Ur || !Dr || !Tr) return fail(DQMC_E_ARG, "null UDT factor");
  Ops o(ctx->b->groups[0]);
  RC(put_mat(ctx->b, NM(M_UL), Ul));
  RC(put_mat(ctx->b, NM(M_TL), Tl));
  RC(put_mat(ctx->b, NM(M_UR), Ur));
  RC(put_mat(ctx->b, NM(M_TR), Tr));
  RC(put_vec(ctx->b, NV(V_DL), Dl));
  RC(put_vec(ctx->b, NV(V_DR), Dr));
  RC(o.calculate_greens(true));
  RC(batch_sync(ctx->b));
  if (G) RC(get_mat(ctx->b, NM(M_G), G));
  if (logdet) *logdet = read_scalar(ctx->b, 0, S_LOGDET);
  return 0;
}

int dqmc_greens_at(dqmc_ctx* ctx, int slice, double* G, double* logdet) {
  // calc_greens_and_logdet (fermion_measurements.jl:1061-1101) via calc_Bchain(_dagger)
  CTX_CHECK(ctx);
  Group& g = ctx->b->groups[0];
  if (slice < 1 || slice > g.M) return fail(DQMC_E_ARG, "slice out of range");
  Ops o(g);
  const int sm = g.s;
  // Ur, Dr, Tr = B(slice)' ... B(M)' (calc_Bchain_dagger); Ul, Dl, Tl = B(slice-1) ... B(1)
  RC(o.bchain(slice, g.M, B_DAG_LEFT, NM(M_UR), NV(V_DR), NM(M_TR)));
  if (slice > 1) {
    RC(o.bchain(1, slice - 1, B_LEFT, NM(M_UL), NV(V_DL), NM(M_TL)));
  } else {
    LCHK(launch_identity(g.stream, g.d_desc, 1, g.n, NM(M_UL), NOVEC));
    LCHK(launch_identity(g.stream, g.d_desc, 1, g.n, NM(M_TL), NOVEC));
    LCHK(launch_ones_vec(g.stream, g.d_desc, 1, g.n, NV(V_DL)));
  }
  // keep the chain's G and log_det: compute into them, then restore
  double* sc = ctx->b->chains[0].d.scal;
  HIPCHK(hipMemcpyAsync(sc + S_LOGDET_SAVE, sc + S_LOGDET, sizeof(double), hipMemcpyDeviceToDevice, g.stream));
  RC(o.copy(NM(M_GW), NM(M_G)));
  RC(o.calculate_greens(true));
  RC(batch_sync(ctx->b));
  if (G) RC(get_mat(ctx->b, NM(M_G), G));
  if (logdet) *logdet = read_scalar(ctx->b, 0, S_LOGDET);
  RC(o.copy(NM(M_G), NM(M_GW)));
  HIPCHK(hipMemcpyAsync(sc + S_LOGDET, sc + S_LOGDET_SAVE, sizeof(double), hipMemcpyDeviceToDevice, g.stream));
  return batch_sync(ctx->b);
}

int dqmc_global_update(dqmc_ctx* ctx, double box_global, int* accepted) {
  // global_update (global_updates.jl:19-58). The reference swaps host stacks into the
  // backup (:1-8); here the backup is a device copy and the restore on rejection is a
  // kernel conditioned on the device-side decision, so nothing waits for the host.
  CTX_CHECK(ctx);
  dqmc_batch* b = ctx->b;
  Group& g = b->groups[0];
  ChainMem& cm = b->chains[0];
  if (g.cur != g.M || g.dir != -1)
    return fail(DQMC_E_STATE, "global_update needs current_slice == M and direction == -1");
  if (!(box_global >= 0.0)) return fail(DQMC_E_ARG, "box_global must be >= 0");
  const size_t nn = (size_t)g.n * g.n;
  const size_t n_st = 2 * nn * cm.nelem;  // doubles of one of u_stack / t_stack
  const size_t n_d = (size_t)g.n * cm.nelem, n_g = 2 * nn, n_phi = (size_t)3 * g.N * g.M;
  if (!cm.gb) HIPCHK(hipMalloc(&cm.gb, sizeof(double) * (2 * n_st + n_d + n_g + n_phi)));
  double* gu = cm.gb;
  double* gt = gu + n_st;
  double* gd = gt + n_st;
  double* gg = gd + n_d;
  double* gphi = gg + n_g;
  hipStream_t st = g.stream;
  ChainDesc& d = cm.d;
  HIPCHK(hipMemcpyAsync(gu, d.ustack, sizeof(double) * n_st, hipMemcpyDeviceToDevice, st));
  HIPCHK(hipMemcpyAsync(gt, d.tstack, sizeof(double) * n_st, hipMemcpyDeviceToDevice, st));
  HIPCHK(hipMemcpyAsync(gd, d.dstack, sizeof(double) * n_d, hipMemcpyDeviceToDevice, st));
  HIPCHK(hipMemcpyAsync(gg, d.mat[M_G], sizeof(double) * n_g, hipMemcpyDeviceToDevice, st));
  HIPCHK(hipMemcpyAsync(d.scal + S_GB_ACTION, d.scal + S_BOSON_ACTION, sizeof(double), hipMemcpyDeviceToDevice,
                        st));
  HIPCHK(hipMemcpyAsync(d.scal + S_GB_LOGDET, d.scal + S_LOGDET, sizeof(double), hipMemcpyDeviceToDevice, st));
  LCHK(launch_global_shift(st, g.d_desc, gphi, box_global));
  Ops o(g);
  RC(o.build_stack());
  RC(o.propagate());
  LCHK(launch_boson_action(st, g.d_desc, 1, g.N));
  LCHK(launch_global_decide(st, g.d_desc));
  RestoreList r;
  r.n = 5;
  r.dst[0] = (double*)d.ustack;
  r.src[0] = gu;
  r.count[0] = (long long)n_st;
  r.dst[1] = (double*)d.tstack;
  r.src[1] = gt;
  r.count[1] = (long long)n_st;
  r.dst[2] = d.dstack;
  r.src[2] = gd;
  r.count[2] = (long long)n_d;
  r.dst[3] = (double*)d.mat[M_G];
  r.src[3] = gg;
  r.count[3] = (long long)n_g;
  r.dst[4] = d.phi;
  r.src[4] = gphi;
  r.count[4] = (long long)n_phi;
  LCHK(launch_global_restore(st, d.scal + S_GLOBAL_ACC, r));
  if (accepted) {
    RC(batch_sync(b));
    *accepted = (int)read_scalar(b, 0, S_GLOBAL_ACC);
  }
  return 0;
}

int dqmc_effective_greens2greens(dqmc_ctx* ctx, double* G) {
  CTX_CHECK(ctx);
  if (!G) return fail(DQMC_E_ARG, "null G");
  Ops o(ctx->b->groups[0]);
  RC(put_mat(ctx->b, NM(M_GRAD), G));
  RC(o.effective_greens(NM(M_GRAD)));
  RC(get_mat(ctx->b, NM(M_GRAD), G));
  return check_device_error(ctx->b);
}

int dqmc_tdgf_alloc(dqmc_ctx* ctx) {
  CTX_CHECK(ctx);
  dqmc_batch* b = ctx->b;
  Group& g = b->groups[0];
  ChainMem& cm = b->chains[0];
  if (cm.tdgf_mat) return 0;
  if (g.M % 2) return fail(DQMC_E_ARG, "time-displaced Green's functions need an even number of slices");
  const size_t nn = (size_t)g.n * g.n;
  const int R = g.M / g.s;
  const size_t mat_bytes = ((size_t)8 * R + 2 * (size_t)g.M) * nn * sizeof(cplx);
  const size_t vec_bytes = (size_t)4 * R * g.n * sizeof(double);
  void *m = nullptr, *v = nullptr;
  hipError_t e = hipMalloc(&m, mat_bytes);
  if (e == hipSuccess) e = hipMalloc(&v, vec_bytes);
  if (e == hipSuccess) e = hipMemset(m, 0, mat_bytes);
  if (e == hipSuccess) e = hipMemset(v, 0, vec_bytes);
  if (e != hipSuccess) {
    if (m) (void)hipFree(m);
    if (v) (void)hipFree(v);
    return fail(DQMC_E_HIP, std::string("time-displaced Green's function storage: ") + hipGetErrorString(e));
  }
  cm.tdgf_mat = (cplx*)m;
  cm.tdgf_vec = (double*)v;
  cm.d.xmat = cm.tdgf_mat;
  cm.d.xvec = cm.tdgf_vec;
  return upload_desc(b, g);
}

int dqmc_tdgf_free(dqmc_ctx* ctx) {
  CTX_CHECK(ctx);
  dqmc_batch* b = ctx->b;
  ChainMem& cm = b->chains[0];
  RC(batch_sync(b));
  if (cm.tdgf_mat) (void)hipFree(cm.tdgf_mat);
  if (cm.tdgf_vec) (void)hipFree(cm.tdgf_vec);
  cm.tdgf_mat = nullptr;
  cm.tdgf_vec = nullptr;
  cm.d.xmat = nullptr;
  cm.d.xvec = nullptr;
  return upload_desc(b, b->groups[0]);
}

int dqmc_measure_tdgfs(dqmc_ctx* ctx) {
  CTX_CHECK(ctx);
  if (!ctx->b->chains[0].tdgf_mat) return fail(DQMC_E_STATE, "dqmc_tdgf_alloc has not been called");
  Ops o(ctx->b->groups[0]);
  RC(o.measure_tdgfs());
  return batch_sync(ctx->b);
}

int dqmc_get_tdgf(dqmc_ctx* ctx, int tau, double* Gt0, double* G0t) {
  CTX_CHECK(ctx);
  Group& g = ctx->b->groups[0];
  if (!ctx->b->chains[0].tdgf_mat) return fail(DQMC_E_STATE, "dqmc_tdgf_alloc has not been called");
  if (tau < 1 || tau > g.M) return fail(DQMC_E_ARG, "tau out of range");
  Ops o(g);
  if (Gt0) RC(get_mat(ctx->b, o.GT0(tau), Gt0));
  if (G0t) RC(get_mat(ctx->b, o.G0T(tau), G0t));
  return 0;
}

int dqmc_get_tdgf_stack(dqmc_ctx* ctx, int which, int k, double* U, double* D, double* T) {
  CTX_CHECK(ctx);
  Group& g = ctx->b->groups[0];
  if (!ctx->b->chains[0].tdgf_mat) return fail(DQMC_E_STATE, "dqmc_tdgf_alloc has not been called");
  Ops o(g);
  if (which < 0 || which > 3 || k < 0 || k >= o.nr()) return fail(DQMC_E_ARG, "bad stack index");
  RC(batch_sync(ctx->b));
  if (U) RC(get_mat(ctx->b, o.XU(which, k), U));
  if (T) RC(get_mat(ctx->b, o.XT(which, k), T));
  if (D) RC(get_vec(ctx->b, o.XD(which, k), D));
  return 0;
}

int dqmc_tdgf_direct(dqmc_ctx* ctx, int slice, double* G) {
  // calc_tdgf_direct (fermion_measurements.jl:1553-1576), scalettar form
  CTX_CHECK(ctx);
  Group& g = ctx->b->groups[0];
  if (!G || slice < 1 || slice > g.M) return fail(DQMC_E_ARG, "bad tdgf_direct args");
  Ops o(g);
  if (slice != 1) {
    RC(o.bchain(1, slice - 1, B_INV_LEFT, NM(M_UL), NV(V_DL), NM(M_TL)));
  } else {
    LCHK(launch_identity(g.stream, g.d_desc, 1, g.n, NM(M_UL), NOVEC));
    LCHK(launch_identity(g.stream, g.d_desc, 1, g.n, NM(M_TL), NOVEC));
    LCHK(launch_ones_vec(g.stream, g.d_desc, 1, g.n, NV(V_DL)));
  }
  if (slice != g.M) {
    RC(o.bchain(slice, g.M, B_LEFT, NM(M_UR), NV(V_DR), NM(M_TR)));
  } else {
    LCHK(launch_identity(g.stream, g.d_desc, 1, g.n, NM(M_UR), NOVEC));
    LCHK(launch_identity(g.stream, g.d_desc, 1, g.n, NM(M_TR), NOVEC));
    LCHK(launch_ones_vec(g.stream, g.d_desc, 1, g.n, NV(V_DR)));
  }
  RC(o.inv_sum_udts_scalettar(MatRef{-1, 0}, NM(M_UL), NV(V_DL), NM(M_TL), NM(M_UR), NV(V_DR), NM(M_TR),
                              NM(M_CURU), NV(V_D6), NM(M_GW)));
  RC(o.effective_greens_ut(NM(M_CURU), NM(M_GW)));
  LCHK(launch_scale_cols(g.stream, g.d_desc, 1, g.n, NM(M_CURU), NV(V_D6)));
  RC(o.gemm(NM(M_GRAD), NM(M_CURU), OP_N, NM(M_GW), OP_N));
  RC(batch_sync(ctx->b));
  return get_mat(ctx->b, NM(M_GRAD), G);
}

int dqmc_detratio(dqmc_ctx* ctx, int site, const double* new_op, double* det, double* delta, double* Mout) {
  CTX_CHECK(ctx);
  dqmc_batch* b = ctx->b;
  Group& g = b->groups[0];
  if (site < 0 || site >= g.N || !new_op) return fail(DQMC_E_ARG, "bad detratio args");
  if (g.cur < 1 || g.cur > g.M) return fail(DQMC_E_STATE, "detratio outside slices 1..M");
  HIPCHK(hipMemcpyAsync(b->d_scratch + 100, new_op, 3 * sizeof(double), hipMemcpyHostToDevice, g.stream));
  LCHK(launch_detratio(g.stream, g.d_desc, g.cur - 1, site, b->d_scratch + 100, b->d_scratch));
  double out[66];
  HIPCHK(hipMemcpyAsync(out, b->d_scratch, sizeof(out), hipMemcpyDeviceToHost, g.stream));
  HIPCHK(hipStreamSynchronize(g.stream));
  if (det) memcpy(det, out, 2 * sizeof(double));
  if (delta) memcpy(delta, out + 2, 32 * sizeof(double));
  if (Mout) memcpy(Mout, out + 34, 32 * sizeof(double));
  return 0;
}

int dqmc_update_greens(dqmc_ctx* ctx, int site) {
  CTX_CHECK(ctx);
  Group& g = ctx->b->groups[0];
  if (site < 0 || site >= g.N) return fail(DQMC_E_ARG, "bad site");
  LCHK(launch_force_update(g.stream, g.d_desc, g.n, site));
  return batch_sync(ctx->b);
}

int dqmc_gemm(dqmc_ctx* ctx, int opA, int opB, const double* A, const double* B, double* C) {
  CTX_CHECK(ctx);
  if (!A || !B || !C || opA < 0 || opA > 1 || opB < 0 || opB > 1) return fail(DQMC_E_ARG, "bad gemm args");
  Ops o(ctx->b->groups[0]);
  RC(put_mat(ctx->b, NM(M_TMP), A));
  RC(put_mat(ctx->b, NM(M_TMP2), B));
  RC(o.gemm(NM(M_TMP3), NM(M_TMP), (GemmOp)opA, NM(M_TMP2), (GemmOp)opB));
  return get_mat(ctx->b, NM(M_TMP3), C);
}

int dqmc_inverse(dqmc_ctx* ctx, const double* A, double* Ainv, double* logabsdet) {
  CTX_CHECK(ctx);
  if (!A || !Ainv) return fail(DQMC_E_ARG, "bad inverse args");
  Ops o(ctx->b->groups[0]);
  RC(put_mat(ctx->b, NM(M_TMP4), A));
  RC(o.inverse(NM(M_TMP4)));
  RC(get_mat(ctx->b, NM(M_TMP4), Ainv));
  RC(check_device_error(ctx->b));
  if (logabsdet) *logabsdet = read_scalar(ctx->b, 0, S_LOGDET_A);
  return 0;
}

int dqmc_boson_action(dqmc_ctx* ctx, double* S) {
  CTX_CHECK(ctx);
  Group& g = ctx->b->groups[0];
  // computed into a scratch copy of the scalar so the running p.boson_action is kept
  double keep = read_scalar(ctx->b, 0, S_BOSON_ACTION);
  LCHK(launch_boson_action(g.stream, g.d_desc, 1, g.N));
  RC(batch_sync(ctx->b));
  if (S) *S = read_scalar(ctx->b, 0, S_BOSON_ACTION);
  HIPCHK(hipMemcpy(ctx->b->chains[0].d.scal + S_BOSON_ACTION, &keep, sizeof(double), hipMemcpyHostToDevice));
  return 0;
}

// dS/dphi of the current field into the context's gradient buffer (enqueue only):
// build_stack, then one down pass over all slices; at slice l the Green's function G(l)
// of the stack (stack.jl:336) gives the fermionic part, the bosonic part is analytic.
// Afterwards scal[S_BOSON_ACTION] = S_B and scal[S_LOGDET] = log|det G|.
static int grad_enqueue(dqmc_batch* b) {
  Group& g = b->groups[0];
  ChainMem& cm = b->chains[0];
  Ops o(g);
  int rc;
  if (!cm.grad) {
    HIPCHK(hipMalloc(&cm.grad, sizeof(double) * 3 * cm.N * cm.M));
    HIPCHK(hipMalloc(&cm.grad_ptr, sizeof(double*)));
    HIPCHK(hipMemcpy(cm.grad_ptr, &cm.grad, sizeof(double*), hipMemcpyHostToDevice));
    HIPCHK(hipMalloc(&cm.mom, sizeof(double) * 3 * cm.N * cm.M));
    HIPCHK(hipMalloc(&cm.phi_bak, sizeof(double) * 3 * cm.N * cm.M));
  }
  if ((rc = o.build_stack())) return rc;
  if ((rc = o.propagate())) return rc;  // -> slice M, direction -1, log det of G(1)
  LCHK(launch_boson_grad(g.stream, g.d_desc, 1, cm.N * cm.M, cm.grad_ptr));
  LCHK(launch_boson_action(g.stream, g.d_desc, 1, g.N));
  for (int step = 0; step < g.M; ++step) {
    LCHK(launch_grad_slice(g.stream, g.d_desc, 1, g.N, g.cur - 1, cm.grad_ptr));
    if (step + 1 < g.M && (rc = o.propagate())) return rc;
  }
  return 0;
}

int dqmc_grad(dqmc_ctx* ctx, double* grad, double* S) {
  CTX_CHECK(ctx);
  dqmc_batch* b = ctx->b;
  ChainMem& cm = b->chains[0];
  RC(grad_enqueue(b));
  RC(batch_sync(b));
  if (grad) HIPCHK(hipMemcpy(grad, cm.grad, sizeof(double) * 3 * cm.N * cm.M, hipMemcpyDeviceToHost));
  if (S) *S = read_scalar(b, 0, S_BOSON_ACTION) + read_scalar(b, 0, S_LOGDET);
  return 0;
}

int dqmc_hmc_trajectory(dqmc_ctx* ctx, const double* momenta, int nsteps, double eps, double u, int* accepted,
                        double* dH) {
  CTX_CHECK(ctx);
  dqmc_batch* b = ctx->b;
  Group& g = b->groups[0];
  ChainMem& cm = b->chains[0];
  if (!momenta || nsteps < 1 || !(eps > 0)) return fail(DQMC_E_ARG, "bad HMC arguments");
  const long long nf = 3LL * cm.N * cm.M;
  double k0 = 0.0;
  for (long long e = 0; e < nf; ++e) k0 += 0.5 * momenta[e] * momenta[e];
  RC(grad_enqueue(b));  // allocates the buffers on first use
  HIPCHK(hipMemcpyAsync(cm.mom, momenta, sizeof(double) * nf, hipMemcpyHostToDevice, g.stream));
  HIPCHK(hipMemcpyAsync(cm.phi_bak, cm.d.phi, sizeof(double) * nf, hipMemcpyDeviceToDevice, g.stream));
  RC(batch_sync(b));
  const double S0 = read_scalar(b, 0, S_BOSON_ACTION) + read_scalar(b, 0, S_LOGDET);
  // leapfrog: p -= eps/2 dS; (phi += eps p; dS; p -= eps dS) ...; last kick eps/2
  LCHK(launch_daxpy(g.stream, cm.mom, cm.grad, -0.5 * eps, nf));
  for (int s = 0; s < nsteps; ++s) {
    LCHK(launch_daxpy(g.stream, cm.d.phi, cm.mom, eps, nf));
    RC(grad_enqueue(b));
    LCHK(launch_daxpy(g.stream, cm.mom, cm.grad, (s + 1 < nsteps) ? -eps : -0.5 * eps, nf));
  }
  RC(batch_sync(b));
  const double S1 = read_scalar(b, 0, S_BOSON_ACTION) + read_scalar(b, 0, S_LOGDET);
  std::vector<double> p1(nf);
  HIPCHK(hipMemcpy(p1.data(), cm.mom, sizeof(double) * nf, hipMemcpyDeviceToHost));
  double k1 = 0.0;
  for (long long e = 0; e < nf; ++e) k1 += 0.5 * p1[e] * p1[e];
  const double dh = (S1 + k1) - (S0 + k0);
  const int acc = (dh <= 0.0) || (u < std::exp(-dh));
  if (!acc) HIPCHK(hipMemcpyAsync(cm.d.phi, cm.phi_bak, sizeof(double) * nf, hipMemcpyDeviceToDevice, g.stream));
  // leave the chain as init! would: boson action, build_stack, propagate
  RC(group_init(b, g));
  RC(batch_sync(b));
  if (accepted) *accepted = acc;
  if (dH) *dH = dh;
  return 0;
}

// ---- fermionic measurements (fermion_measurements.jl:216-229, 363-957) --------------------
}  // extern "C"
namespace {

int meas_alloc(ChainMem& cm) {
  if (cm.meas) return 0;
  const size_t nn = (size_t)cm.n * cm.n, N = (size_t)cm.N;
  const size_t blocks = (size_t)std::max(fermion_obs_blocks(cm.M, cm.N), fermion_obs_blocks(1, cm.N));
  const size_t tot = nn + 16 * N + blocks * 2 * N + 2 * N;
  void* p = nullptr;
  HIPCHK(hipMalloc(&p, tot * sizeof(cplx)));
  cm.meas = (cplx*)p;
  cm.meas_etgf = cm.meas;
  cm.meas_tnb = cm.meas_etgf + nn;
  cm.meas_part = cm.meas_tnb + 16 * N;
  cm.meas_res = cm.meas_part + blocks * 2 * N;
  return 0;
}

// one observable family on the device, the two L x L result arrays to the host (complex, or
// only the real parts when `real_only`)
int meas_run(dqmc_ctx* ctx, int kind, bool zf, double* out0, double* out1, bool real_only) {
  dqmc_batch* b = ctx->b;
  ChainMem& cm = b->chains[0];
  Group& g = b->groups[0];
  if (!cm.has_etgf && !(kind == 0 && zf))
    return fail(DQMC_E_STATE, "no equal-time Green's function: call dqmc_measure_greens or dqmc_set_etgf first");
  if (zf && !cm.tdgf_mat) return fail(DQMC_E_STATE, "zero-frequency observables need dqmc_tdgf_alloc + dqmc_measure_tdgfs");
  if (kind == 2 && !cm.has_hopping) return fail(DQMC_E_STATE, "current correlations need dqmc_set_hopping_matrix");
  int rc = meas_alloc(cm);
  if (rc) return rc;
  const size_t nn = (size_t)cm.n * cm.n;
  const cplx* E = cm.meas_etgf;
  if (!zf) {
    LCHK(launch_fermion_obs(g.stream, kind, cm.L, 1, E, E, E, 0, E, 0, cm.meas_tnb, cm.meas_part, cm.meas_res));
  } else {
    Ops o(g);
    const cplx* gt0 = cm.tdgf_mat + nn * (size_t)o.GT0(1).idx;
    const cplx* g0t = cm.tdgf_mat + nn * (size_t)o.G0T(1).idx;
    LCHK(launch_fermion_obs(g.stream, kind, cm.L, cm.M, E, E, gt0, nn, g0t, nn, cm.meas_tnb, cm.meas_part,
                            cm.meas_res));
  }
  std::vector<cplx> res((size_t)2 * cm.N);
  HIPCHK(hipMemcpyAsync(res.data(), cm.meas_res, res.size() * sizeof(cplx), hipMemcpyDeviceToHost, g.stream));
  HIPCHK(hipStreamSynchronize(g.stream));
  RC(check_device_error(b));
  double* outs[2] = {out0, out1};
  for (int k = 0; k < 2; ++k) {
    if (!outs[k]) continue;
    for (int r = 0; r < cm.N; ++r) {
      const cplx v = res[(size_t)k * cm.N + r];
      if (real_only) {
        outs[k][r] = v.re;
      } else {
        outs[k][2 * r] = v.re;
        outs[k][2 * r + 1] = v.im;
      }
    }
  }
  return 0;
}

}  // namespace
extern "C" {

int dqmc_set_hopping_matrix(dqmc_ctx* ctx, const double* T) {
  CTX_CHECK(ctx);
  if (!T) return fail(DQMC_E_ARG, "null hopping matrix");
  ChainMem& cm = ctx->b->chains[0];
  RC(batch_sync(ctx->b));
  RC(meas_alloc(cm));
  const int L = cm.L, N = cm.N, n = cm.n;
  std::vector<double> h((size_t)2 * 16 * N);
  for (int d = 0; d < 2; ++d)
    for (int f = 0; f < 4; ++f)
      for (int i = 0; i < N; ++i) {
        const int x = i / L, y = i - x * L;
        const int j = d == 0 ? y + L * ((x + 1) % L) : (y + 1) % L + L * x;
        const size_t a = (size_t)(f * N + j) * n + (f * N + i);  // T[fN + i, fN + j]
        const size_t c = (size_t)(f * N + i) * n + (f * N + j);  // T[fN + j, fN + i]
        double* o = h.data() + 4 * ((size_t)(d * 4 + f) * N + i);
        o[0] = T[2 * a];
        o[1] = T[2 * a + 1];
        o[2] = T[2 * c];
        o[3] = T[2 * c + 1];
      }
  HIPCHK(hipMemcpy(cm.meas_tnb, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice));
  cm.has_hopping = true;
  return 0;
}

int dqmc_measure_greens(dqmc_ctx* ctx, double* G) {
  // measure_greens (fermion_measurements.jl:225-229): calc_greens(mc, 1) (from scratch, as
  // calc_greens_and_logdet :1061-1101), then effective_greens2greens! (:1125-1141)
  CTX_CHECK(ctx);
  dqmc_batch* b = ctx->b;
  Group& g = b->groups[0];
  ChainMem& cm = b->chains[0];
  RC(meas_alloc(cm));
  Ops o(g);
  RC(o.bchain(1, g.M, B_DAG_LEFT, NM(M_UR), NV(V_DR), NM(M_TR)));
  LCHK(launch_identity(g.stream, g.d_desc, 1, g.n, NM(M_UL), NOVEC));
  LCHK(launch_identity(g.stream, g.d_desc, 1, g.n, NM(M_TL), NOVEC));
  LCHK(launch_ones_vec(g.stream, g.d_desc, 1, g.n, NV(V_DL)));
  double* sc = cm.d.scal;
  HIPCHK(hipMemcpyAsync(sc + S_LOGDET_SAVE, sc + S_LOGDET, sizeof(double), hipMemcpyDeviceToDevice, g.stream));
  RC(o.copy(NM(M_GW), NM(M_G)));
  RC(o.calculate_greens(true));
  RC(o.effective_greens(NM(M_G)));
  HIPCHK(hipMemcpyAsync(cm.meas_etgf, cm.d.mat[M_G], sizeof(cplx) * cm.n * cm.n, hipMemcpyDeviceToDevice, g.stream));
  RC(o.copy(NM(M_G), NM(M_GW)));
  HIPCHK(hipMemcpyAsync(sc + S_LOGDET, sc + S_LOGDET_SAVE, sizeof(double), hipMemcpyDeviceToDevice, g.stream));
  cm.has_etgf = true;
  if (G) HIPCHK(hipMemcpyAsync(G, cm.meas_etgf, sizeof(cplx) * cm.n * cm.n, hipMemcpyDeviceToHost, g.stream));
  return batch_sync(b);
}

int dqmc_measure_current_greens(dqmc_ctx* ctx, double* G) {
  // _measure_observables! (dqmc_framework.jl:378-382): wrap_greens(mc, mc.s.greens,
  // current_slice, 1), then effective_greens2greens!; the chain's G is left unchanged
  CTX_CHECK(ctx);
  dqmc_batch* b = ctx->b;
  Group& g = b->groups[0];
  ChainMem& cm = b->chains[0];
  if (g.cur < 1 || g.cur > g.M) return fail(DQMC_E_STATE, "the chain is not at a slice 1..M");
  RC(meas_alloc(cm));
  Ops o(g);
  RC(o.copy(NM(M_GRAD), NM(M_G)));
  RC(o.wrap(NM(M_GRAD), g.cur, 1));
  RC(o.effective_greens(NM(M_GRAD)));
  HIPCHK(hipMemcpyAsync(cm.meas_etgf, cm.d.mat[M_GRAD], sizeof(cplx) * cm.n * cm.n, hipMemcpyDeviceToDevice,
                        g.stream));
  cm.has_etgf = true;
  if (G) HIPCHK(hipMemcpyAsync(G, cm.meas_etgf, sizeof(cplx) * cm.n * cm.n, hipMemcpyDeviceToHost, g.stream));
  return batch_sync(b);
}

int dqmc_set_etgf(dqmc_ctx* ctx, const double* G) {
  CTX_CHECK(ctx);
  if (!G) return fail(DQMC_E_ARG, "null G");
  ChainMem& cm = ctx->b->chains[0];
  RC(batch_sync(ctx->b));
  RC(meas_alloc(cm));
  HIPCHK(hipMemcpy(cm.meas_etgf, G, sizeof(cplx) * cm.n * cm.n, hipMemcpyHostToDevice));
  cm.has_etgf = true;
  return 0;
}

int dqmc_measure_etpc(dqmc_ctx* ctx, double* plus, double* minus) {
  CTX_CHECK(ctx);
  return meas_run(ctx, 0, false, minus, plus, true);
}
int dqmc_measure_zfpc(dqmc_ctx* ctx, double* plus, double* minus) {
  CTX_CHECK(ctx);
  return meas_run(ctx, 0, true, minus, plus, true);
}
int dqmc_measure_etcdc(dqmc_ctx* ctx, double* plus, double* minus) {
  CTX_CHECK(ctx);
  return meas_run(ctx, 1, false, minus, plus, false);
}
int dqmc_measure_zfcdc(dqmc_ctx* ctx, double* plus, double* minus) {
  CTX_CHECK(ctx);
  return meas_run(ctx, 1, true, minus, plus, false);
}
int dqmc_measure_zfccc(dqmc_ctx* ctx, double* xx, double* yy, double* sfdensity) {
  CTX_CHECK(ctx);
  ChainMem& cm = ctx->b->chains[0];
  const int L = cm.L;
  std::vector<double> bx((size_t)2 * cm.N), by((size_t)2 * cm.N);
  RC(meas_run(ctx, 2, true, bx.data(), by.data(), false));
  if (xx) memcpy(xx, bx.data(), bx.size() * sizeof(double));
  if (yy) memcpy(yy, by.data(), by.size() * sizeof(double));
  if (sfdensity) {
    // measure_sfdensity (:911-923): rfft of real(zfccc_xx); 1/4 (Re q=(ky 0, kx 1) - Re q=(ky 1, kx 0))
    const double w = 2.0 * M_PI / L;
    double qx = 0.0, qy = 0.0;
    for (int x = 0; x < L; ++x)
      for (int y = 0; y < L; ++y) {
        const double v = bx[2 * ((size_t)y + (size_t)L * x)];
        qx += v * std::cos(w * x);
        qy += v * std::cos(w * y);
      }
    *sfdensity = 0.25 * (qx - qy);
  }
  return 0;
}

// ---- batches ---------------------------------------------------------------
int dqmc_batch_create(const dqmc_params* p, const dqmc_operators* ops, int nchains, int device, dqmc_batch** out) {
  return batch_create(p, ops, nchains, device, out);
}
int dqmc_batch_destroy(dqmc_batch* b) { return batch_destroy(b); }

int dqmc_batch_set_field(dqmc_batch* b, int chain, const double* phi) {
  if (!b || !phi || chain < 0 || chain >= (int)b->chains.size()) return fail(DQMC_E_ARG, "bad set_field");
  ChainMem& cm = b->chains[chain];
  RC(batch_sync(b));
  HIPCHK(hipMemcpy(cm.d.phi, phi, sizeof(double) * 3 * cm.N * cm.M, hipMemcpyHostToDevice));
  return 0;
}
int dqmc_batch_get_field(dqmc_batch* b, int chain, double* phi) {
  if (!b || !phi || chain < 0 || chain >= (int)b->chains.size()) return fail(DQMC_E_ARG, "bad get_field");
  ChainMem& cm = b->chains[chain];
  RC(batch_sync(b));
  HIPCHK(hipMemcpy(phi, cm.d.phi, sizeof(double) * 3 * cm.N * cm.M, hipMemcpyDeviceToHost));
  return 0;
}
int dqmc_batch_set_uniforms(dqmc_batch* b, int chain, const double* u, long long count) {
  if (!b) return fail(DQMC_E_ARG, "null batch");
  return set_uniforms(b, chain, u, count);
}
int dqmc_batch_init(dqmc_batch* b) {
  if (!b) return fail(DQMC_E_ARG, "null batch");
  for (auto& g : b->groups) RC(group_init(b, g));
  return 0;
}
int dqmc_batch_sweep(dqmc_batch* b, int nsweeps, const double* box) {
  if (!b || !box || nsweeps < 0) return fail(DQMC_E_ARG, "bad sweep args");
  // groups are independent streams: enqueue them step-interleaved so their kernels overlap
  for (auto& g : b->groups) {
    const double bx = box[g.chains[0]];
    for (int ci : g.chains)
      if (box[ci] != bx) return fail(DQMC_E_ARG, "chains of one shape group must share the box size");
  }
  if (b->merged && b->groups.size() > 1) {
    bool ok = true;
    const Group& g0 = b->groups[0];
    for (auto& g : b->groups)
      ok = ok && g.s == g0.s && g.chkr == g0.chkr && g.ngroups == g0.ngroups && g.all_checks == g0.all_checks &&
           g.greens_route == 0 && g.udt_route == g0.udt_route && g.local_variant == g0.local_variant &&
           box[g.chains[0]] == box[g0.chains[0]] && g.cur >= 1 && g.cur <= g.M && g.s <= 16;
    if (ok) {
      int dolog = 0;
      for (auto& cm : b->chains) dolog |= cm.log_cap > 0;
      bool graphs = b->use_graph;
      for (auto& g : b->groups) graphs = graphs && g.prof_mask == 0;  // timing events stay eager
      if (graphs) {
        for (int sw = 0; sw < nsweeps; ++sw) RC(merged_sweep_graph(b, box[g0.chains[0]], dolog));
        return 0;
      }
      return merged_sweeps(b, nsweeps, box[g0.chains[0]], dolog);
    }
  }
  bool graphs = b->use_graph;
  for (auto& g : b->groups) graphs = graphs && g.prof_mask == 0;  // timing events stay eager
  if (graphs) {
    for (int sw = 0; sw < nsweeps; ++sw)
      for (auto& g : b->groups) {
        int dolog = 0;
        for (int ci : g.chains) dolog |= b->chains[ci].log_cap > 0;
        RC(group_sweep_graph(g, box[g.chains[0]], dolog));
      }
    return 0;
  }
  std::vector<int> steps(b->groups.size(), 0);
  bool more = true;
  while (more) {
    more = false;
    for (size_t gi = 0; gi < b->groups.size(); ++gi) {
      Group& g = b->groups[gi];
      const int total = nsweeps * g.M;
      if (steps[gi] >= total) continue;
      Ops o(g);
      int dolog = 0;
      for (int ci : g.chains) dolog |= b->chains[ci].log_cap > 0;
      RC(o.propagate());
      RC(o.local_updates(box[g.chains[0]], dolog));
      steps[gi] += 1;
      more = more || steps[gi] < total;
    }
  }
  return 0;
}
int dqmc_batch_set_merged(dqmc_batch* b, int enable) {
  if (!b) return fail(DQMC_E_ARG, "null batch");
  RC(batch_sync(b));
  b->merged = enable != 0;
  return 0;
}
int dqmc_batch_synchronize(dqmc_batch* b) {
  if (!b) return fail(DQMC_E_ARG, "null batch");
  return batch_sync(b);
}
int dqmc_batch_get_greens(dqmc_batch* b, int chain, double* G) {
  if (!b || !G || chain < 0 || chain >= (int)b->chains.size()) return fail(DQMC_E_ARG, "bad get_greens");
  RC(batch_sync(b));
  ChainMem& cm = b->chains[chain];
  HIPCHK(hipMemcpy(G, cm.d.mat[M_G], sizeof(cplx) * cm.n * cm.n, hipMemcpyDeviceToHost));
  return 0;
}
int dqmc_batch_get_stats(dqmc_batch* b, int chain, dqmc_stats* st) {
  if (!b) return fail(DQMC_E_ARG, "null batch");
  return get_stats(b, chain, st);
}
int dqmc_batch_profile(dqmc_batch* b, unsigned class_mask) {
  if (!b) return fail(DQMC_E_ARG, "null batch");
  for (auto& g : b->groups) g.prof_mask = class_mask;
  return 0;
}

int dqmc_batch_profile_read(dqmc_batch* b, double* ms, long long* count, int nclass) {
  if (!b || nclass < 0) return fail(DQMC_E_ARG, "bad profile args");
  RC(batch_sync(b));
  for (int k = 0; k < nclass; ++k) {
    if (ms) ms[k] = 0.0;
    if (count) count[k] = 0;
  }
  for (auto& g : b->groups) {
    for (auto& r : g.prof) {
      float t = 0.f;
      HIPCHK(hipEventElapsedTime(&t, r.a, r.b));
      if (r.cls < nclass) {
        if (ms) ms[r.cls] += t;
        if (count) count[r.cls] += 1;
      }
      g.event_pool.push_back(r.a);
      g.event_pool.push_back(r.b);
    }
    g.prof.clear();
  }
  return 0;
}

int dqmc_profile(dqmc_ctx* ctx, unsigned class_mask) {
  CTX_CHECK(ctx);
  return dqmc_batch_profile(ctx->b, class_mask);
}

int dqmc_profile_read(dqmc_ctx* ctx, double* ms, long long* count, int nclass) {
  CTX_CHECK(ctx);
  return dqmc_batch_profile_read(ctx->b, ms, count, nclass);
}

// ---------------------------------------------------------------------------
// Autotuner: per shape group, time the local-update kernel over (variant, workgroups per
// chain) and the blocked QR over (workgroups per chain, panel width) on the group's live
// state, keep the fastest. The chains' Green's function, field, scalars and uniform cursor
// are saved before and restored after every timed launch (the local updates run on
// synthetic uniforms), so tuning leaves the Monte Carlo state bit for bit as it was.
// ---------------------------------------------------------------------------
namespace {

struct TuneSave {
  cplx* G = nullptr;
  double* phi = nullptr;
  double* scal = nullptr;
  long long* cur = nullptr;
  int* err = nullptr;
  double* uni = nullptr;
};

int autotune_group(dqmc_batch* b, Group& g, double box, int reps, double* rep) {
  const int nch = g.nch();
  hipStream_t st = g.stream;
  std::vector<void*> tmp;
  auto talloc = [&](void** p, size_t bytes) -> int {
    if (hipMalloc(p, bytes) != hipSuccess) return fail(DQMC_E_HIP, "autotune: hipMalloc");
    tmp.push_back(*p);
    return 0;
  };
  auto release = [&]() {
    (void)hipStreamSynchronize(st);
    for (void* p : tmp) (void)hipFree(p);
    tmp.clear();
  };
  hipEvent_t e0 = nullptr, e1 = nullptr;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  int rc = 0;
  std::vector<TuneSave> sv(nch);
  std::vector<ChainDesc> keep(nch);
  // synthetic uniforms for one slice of proposals (host MT19937, fixed seed)
  std::mt19937_64 eng(20240611);
  std::uniform_real_distribution<double> U(0.0, 1.0);
  for (int k = 0; k < nch && !rc; ++k) {
    ChainMem& cm = b->chains[g.chains[k]];
    const size_t nn = (size_t)cm.n * cm.n, nf = (size_t)3 * cm.N * cm.M;
    if ((rc = talloc((void**)&sv[k].G, nn * sizeof(cplx)))) break;
    if ((rc = talloc((void**)&sv[k].phi, nf * sizeof(double)))) break;
    if ((rc = talloc((void**)&sv[k].scal, kNumScalar * sizeof(double)))) break;
    if ((rc = talloc((void**)&sv[k].cur, sizeof(long long)))) break;
    if ((rc = talloc((void**)&sv[k].err, sizeof(int)))) break;
    if ((rc = talloc((void**)&sv[k].uni, (size_t)4 * cm.N * sizeof(double)))) break;
    std::vector<double> hu((size_t)4 * cm.N);
    for (auto& x : hu) x = U(eng);
    HIPCHK(hipMemcpy(sv[k].uni, hu.data(), hu.size() * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpyAsync(sv[k].G, cm.d.mat[M_G], nn * sizeof(cplx), hipMemcpyDeviceToDevice, st));
    HIPCHK(hipMemcpyAsync(sv[k].phi, cm.d.phi, nf * sizeof(double), hipMemcpyDeviceToDevice, st));
    HIPCHK(hipMemcpyAsync(sv[k].scal, cm.d.scal, kNumScalar * sizeof(double), hipMemcpyDeviceToDevice, st));
    HIPCHK(hipMemcpyAsync(sv[k].cur, cm.d.ucursor, sizeof(long long), hipMemcpyDeviceToDevice, st));
    HIPCHK(hipMemcpyAsync(sv[k].err, cm.d.error, sizeof(int), hipMemcpyDeviceToDevice, st));
    keep[k] = cm.d;
  }
  if (rc) {
    release();
    return rc;
  }
  HIPCHK(hipStreamSynchronize(st));
  auto restore = [&]() -> int {
    for (int k = 0; k < nch; ++k) {
      ChainMem& cm = b->chains[g.chains[k]];
      const size_t nn = (size_t)cm.n * cm.n, nf = (size_t)3 * cm.N * cm.M;
      HIPCHK(hipMemcpyAsync(cm.d.mat[M_G], sv[k].G, nn * sizeof(cplx), hipMemcpyDeviceToDevice, st));
      HIPCHK(hipMemcpyAsync(cm.d.phi, sv[k].phi, nf * sizeof(double), hipMemcpyDeviceToDevice, st));
      HIPCHK(hipMemcpyAsync(cm.d.scal, sv[k].scal, kNumScalar * sizeof(double), hipMemcpyDeviceToDevice, st));
      HIPCHK(hipMemcpyAsync(cm.d.ucursor, sv[k].cur, sizeof(long long), hipMemcpyDeviceToDevice, st));
      HIPCHK(hipMemcpyAsync(cm.d.error, sv[k].err, sizeof(int), hipMemcpyDeviceToDevice, st));
    }
    return 0;
  };
  // the local updates read the synthetic uniforms from a cursor at 0
  long long zero = 0;
  for (int k = 0; k < nch; ++k) {
    ChainMem& cm = b->chains[g.chains[k]];
    cm.d.uniforms = sv[k].uni;
    cm.d.n_uniforms = 4 * cm.N;
  }
  if ((rc = upload_desc(b, g))) {
    release();
    return rc;
  }
  auto timed = [&](auto&& launch, auto&& reset, float* best) -> hipError_t {
    *best = 1e30f;
    for (int r = 0; r <= reps; ++r) {  // the first launch is a warm-up
      if (reset()) return hipErrorUnknown;
      hipError_t e = hipEventRecord(e0, st);
      if (e == hipSuccess) e = launch();
      if (e == hipSuccess) e = hipEventRecord(e1, st);
      if (e == hipSuccess) e = hipEventSynchronize(e1);
      if (e != hipSuccess) {
        (void)hipGetLastError();
        (void)hipStreamSynchronize(st);
        return e;
      }
      float ms = 0.f;
      (void)hipEventElapsedTime(&ms, e0, e1);
      if (r > 0) *best = std::min(*best, ms);
    }
    return hipSuccess;
  };
  // a lone shape group may use several workgroups per CU (the launch fails when they do not
  // fit, and that candidate is skipped); the groups of a mixed batch run concurrently, so
  // their cooperative grids together must fit one workgroup per CU
  const int total = (int)b->chains.size();
  const int cap = b->groups.size() == 1 ? std::max(1, std::min({kMaxWG, std::max(1, 1024 / nch), g.n / 2}))
                                        : std::max(1, std::min({kMaxWG, 256 / total, g.n / 2}));
  std::vector<int> wgs;
  for (int w : {1, 2, 3, 4, 6, 8, 12, 16, 25, 32, 50, 64, 100, 128})
    if (w <= cap) wgs.push_back(w);
  // ---- local updates
  auto reset_local = [&]() -> int {
    if (restore()) return 1;
    for (int k = 0; k < nch; ++k)
      if (hipMemcpyAsync(b->chains[g.chains[k]].d.ucursor, &zero, sizeof(long long), hipMemcpyHostToDevice, st) !=
          hipSuccess)
        return 1;
    return 0;
  };
  float best_local = 1e30f;
  int best_w = g.wg_local, best_v = g.local_variant;
  for (int w : wgs)
    for (int v : {4, 2, 3}) {
      if (v == 4 && local_window_sites(g.n, w) < 2) continue;
      if (v == 2 && !local_tile_fits(g.n, w)) continue;
      if (v == 3 && !local_delayed_fits(g.n)) continue;
      float ms = 0.f;
      const hipError_t e = timed([&]() { return launch_local_updates(st, g.d_desc, nch, g.n, g.cur - 1, nullptr, box,
                                                                     w, 0, v); },
                                 reset_local, &ms);
      if (e == hipSuccess && ms < best_local) {
        best_local = ms;
        best_w = w;
        best_v = v;
      }
    }
  // ---- blocked QR of a copy of G (the stack's factorisation: pre-pivoted)
  auto reset_qr = [&]() -> int {
    for (int k = 0; k < nch; ++k) {
      ChainMem& cm = b->chains[g.chains[k]];
      if (hipMemcpyAsync(cm.d.mat[M_TMP], sv[k].G, (size_t)cm.n * cm.n * sizeof(cplx), hipMemcpyDeviceToDevice, st) !=
          hipSuccess)
        return 1;
    }
    return 0;
  };
  float best_qr = 1e30f;
  int best_wq = g.wg_qr, best_nb = g.qr_nb;
  for (int nb : {16, 8}) {
    if (nb == 16 && g.n > 512) continue;
    for (int w : wgs) {
      if (w > g.n / nb) continue;
      float ms = 0.f;
      const hipError_t e = timed([&]() { return launch_qr_blocked(st, g.d_desc, nch, g.n, MatRef{MK_NAMED, M_TMP},
                                                                  VecRef{VK_NAMED, V_D5}, 1, w, nb); },
                                 reset_qr, &ms);
      if (e == hipSuccess && ms < best_qr) {
        best_qr = ms;
        best_wq = w;
        best_nb = nb;
      }
    }
  }
  // ---- restore the chains and apply
  for (int k = 0; k < nch; ++k) b->chains[g.chains[k]].d = keep[k];
  rc = restore();
  if (!rc) rc = upload_desc(b, g);
  release();
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) return rc;
  g.drop_graphs();
  b->drop_mgraphs();
  if (best_local < 1e29f) {
    g.wg_local = best_w;
    g.local_variant = best_v;
  }
  if (best_qr < 1e29f) {
    g.wg_qr = best_wq;
    g.qr_nb = best_nb;
  }
  if (rep) {
    rep[0] = g.wg_local;
    rep[1] = g.local_variant;
    rep[2] = best_local < 1e29f ? best_local : -1.0;
    rep[3] = g.wg_qr;
    rep[4] = g.qr_nb;
    rep[5] = best_qr < 1e29f ? best_qr : -1.0;
  }
  return 0;
}

}  // namespace

int dqmc_batch_autotune(dqmc_batch* b, const double* box, int reps, double* report) {
  if (!b || !box) return fail(DQMC_E_ARG, "bad autotune args");
  RC(batch_sync(b));
  for (size_t gi = 0; gi < b->groups.size(); ++gi) {
    Group& g = b->groups[gi];
    if (g.cur < 1 || g.cur > g.M) return fail(DQMC_E_STATE, "autotune needs an initialised batch (dqmc_batch_init)");
    RC(autotune_group(b, g, box[g.chains[0]], std::max(1, reps), report ? report + 6 * gi : nullptr));
  }
  return 0;
}

int dqmc_autotune(dqmc_ctx* ctx, double box, int reps, double* report) {
  CTX_CHECK(ctx);
  return dqmc_batch_autotune(ctx->b, &box, reps, report);
}

int dqmc_batch_use_graph(dqmc_batch* b, int enable) {
  if (!b) return fail(DQMC_E_ARG, "null batch");
  RC(batch_sync(b));
  b->use_graph = enable != 0;
  if (!b->use_graph) {
    for (auto& g : b->groups) g.drop_graphs();
    b->drop_mgraphs();
  }
  return 0;
}

int dqmc_use_graph(dqmc_ctx* ctx, int enable) {
  CTX_CHECK(ctx);
  return dqmc_batch_use_graph(ctx->b, enable);
}

}  // extern "C"

// ===========================================================================
// dqmc_inverse_noalloc: dqmc_inverse on caller-provided device buffers, for callers that
// cannot allocate (code under HIP graph capture, serving loops). Its device-side scratch --
// the kernel's descriptor, the pivot-column broadcast buffer, the group-barrier words, the
// error word and the scalar slots -- lives in the fixed array `scratch` below, one copy per
// device for the life of the process. Every call on a device shares that device's copy, so
// calls that overlap on one device race; they must be ordered (one stream, or events).
// ===========================================================================
namespace {

constexpr int kNoAllocMaxN = 1024;
constexpr size_t kNoAllocDescBytes = (sizeof(ChainDesc) + 255) / 256 * 256;
constexpr size_t kNoAllocScalOff = kNoAllocDescBytes;                             // kNumScalar doubles
constexpr size_t kNoAllocBarOff = kNoAllocScalOff + 256;                          // count, gen, error
constexpr size_t kNoAllocBcastOff = kNoAllocBarOff + 256;                         // [2][n] complex
constexpr size_t kNoAllocBytes = kNoAllocBcastOff + (size_t)2 * kNoAllocMaxN * sizeof(cplx);
static_assert(kNumScalar * sizeof(double) <= 256, "scalar slots exceed their scratch slot");

}  // namespace

static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

namespace {

std::mutex g_scratch_mu;
std::map<int, unsigned char*> g_scratch_ptr;  // device -> device address of `scratch`

// The device address of this device's `scratch`, looked up once per device.
int noalloc_scratch(unsigned char** out) {
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  auto it = g_scratch_ptr.find(dev);
  if (it == g_scratch_ptr.end()) {
    void* p = nullptr;
    HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    it = g_scratch_ptr.emplace(dev, (unsigned char*)p).first;
  }
  *out = it->second;
  return 0;
}

// Builds the inverse kernel's descriptor inside the scratch: X (matrix M_TMP4) is the caller's
// A, the gathered inverse (M_INVT) the caller's Ainv; clears the barrier and error words.
__global__ void k_noalloc_inverse_setup(unsigned char* base, int n, cplx* A, cplx* Ainv) {
  if (threadIdx.x != 0) return;
  ChainDesc* d = reinterpret_cast<ChainDesc*>(base);
  *d = ChainDesc{};
  d->n = n;
  d->N = n / 4;
  d->mat[M_TMP4] = A;
  d->mat[M_INVT] = Ainv;
  d->scal = reinterpret_cast<double*>(base + kNoAllocScalOff);
  unsigned int* bar = reinterpret_cast<unsigned int*>(base + kNoAllocBarOff);
  bar[0] = 0u;
  bar[1] = 0u;
  d->bar_count = bar;
  d->bar_gen = bar + 1;
  d->error = reinterpret_cast<int*>(bar + 2);
  *d->error = 0;
  d->inv_bcast = reinterpret_cast<cplx*>(base + kNoAllocBcastOff);
  for (int k = 0; k < kNumScalar; ++k) d->scal[k] = 0.0;
}

// log|det A| to the caller (NaN after a device-side failure, a barrier timeout).
__global__ void k_noalloc_inverse_finish(const unsigned char* base, double* logabsdet) {
  if (threadIdx.x != 0) return;
  const ChainDesc* d = reinterpret_cast<const ChainDesc*>(base);
  *logabsdet = (*d->error != 0) ? __builtin_nan("") : d->scal[S_LOGDET_A];
}

}  // namespace

extern "C" int dqmc_inverse_noalloc(int n, double* A, double* Ainv, double* logabsdet, void* stream) {
  if (n < 1 || n > kNoAllocMaxN) return fail(DQMC_E_ARG, "dqmc_inverse_noalloc: n must be in 1..1024");
  if (!A || !Ainv || A == Ainv) return fail(DQMC_E_ARG, "dqmc_inverse_noalloc: A and Ainv must be distinct buffers");
  unsigned char* base = nullptr;
  int rc = noalloc_scratch(&base);
  if (rc) return rc;
  hipStream_t s = (hipStream_t)stream;
  const int wg = std::max((n + 511) / 512, std::min(std::max(1, n / 8), kMaxWG));
  hipLaunchKernelGGL(k_noalloc_inverse_setup, dim3(1), dim3(64), 0, s, base, n, (cplx*)A, (cplx*)Ainv);
  HIPCHK(hipGetLastError());
  const ChainDesc* desc = reinterpret_cast<const ChainDesc*>(base);
  LCHK(launch_inverse_gather(s, desc, 1, n, NM(M_TMP4), wg));
  if (logabsdet) {
    hipLaunchKernelGGL(k_noalloc_inverse_finish, dim3(1), dim3(64), 0, s, (const unsigned char*)base, logabsdet);
    HIPCHK(hipGetLastError());
  }
  return 0;
}
