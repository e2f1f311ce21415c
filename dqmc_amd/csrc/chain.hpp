// Per-chain device state and the launch wrappers every kernel file exports.
//
// A "batch" is an array of ChainDesc in device memory; kernels index it with
// blockIdx.y so one launch advances every chain of the batch. A single chain
// is a batch of one. Matrices are n x n complex128, column-major (Julia
// ComplexF64 layout), leading dimension n.
#pragma once
#include <vector>
#include "device_common.hpp"

namespace dqmc {

// largest number of bond groups of the generic checkerboard
constexpr int kMaxBondGroups = 16;

// Named n x n complex work matrices of a chain.
enum MatId : int {
  M_G = 0,     // equal-time Green's function (mc.s.greens)
  M_GW,        // wrapped copy for the drift check (mc.s.greens_temp)
  M_GALT,      // ping-pong partner of G inside the local-update kernel
  M_UL, M_TL,  // left UDT operand of calculate_greens
  M_UR, M_TR,  // right UDT operand
  M_TMP, M_TMP2, M_TMP3, M_TMP4,
  M_CURU,      // curr_U of add_slice_sequence_*
  M_QRA,       // factorization workspace (R and reflectors)
  M_QRQ,       // explicit Q of the last factorization
  M_QRT,       // T of the last factorization
  M_UG, M_TG,  // s.U and s.T kept after calculate_greens (logdet)
  M_IDENT,     // identity
  M_GRAD,      // spare n x n for the gradient/TDGF paths
  M_INVT,      // output of the Gauss-Jordan column permutation
  M_ET2,       // dense exp(-dtau T/2)^2        (no checkerboard; operator, read-only)
  M_ETINV2,    // dense exp(+dtau T/2)^2
  M_ET,        // dense exp(-dtau T/2)          (no checkerboard; effective -> true G)
  M_ETINV,     // dense exp(+dtau T/2)
  kNumMat
};
// Named length-n real vectors.
enum VecId : int { V_DL = 0, V_DR, V_D1, V_D2, V_D3, V_D4, V_D5, V_D6, V_DG, V_ONES, V_QRD, kNumVec };

// MK_XMAT / VK_XVEC address the time-displaced Green's function storage (ChainDesc::xmat).
// MK_PLAN / VK_PLAN look the reference up in the chain's own plan table (ChainDesc::plan,
// entry idx): chains advanced together in one launch may then address different stack
// elements or named matrices, or none (entry -1: the kernel skips that chain).
enum MatKind : int { MK_NAMED = 0, MK_USTACK = 1, MK_TSTACK = 2, MK_XMAT = 3, MK_PLAN = 4 };
struct MatRef {
  int kind;
  int idx;
};
enum VecKind : int { VK_NAMED = 0, VK_DSTACK = 1, VK_XVEC = 2, VK_PLAN = 4 };
struct VecRef {
  int kind;
  int idx;
};

struct ChainDesc {
  // model
  int L, N, n, M, safe_mult, nelem, chkr;  // chkr: 1 = Assaad checkerboard, 2 = generic, 0 = dense hopping
  int edrun;
  int plaq_uniform;  // 1: every plaquette of a (variant, group, block) has the same 4x4 block
  double dtau, lam, r, c, u;
  // operators (host-computed at setup, read-only)
  const int* plaq_sites;  // [2 groups][N/4][4] site indices
  const cplx* plaq;       // [4 variants][2 groups][4 blocks][N/4][16] row-major 4x4
  const double* expmu;    // [n] exp(+dtau mu)   (chkr_mu diagonal)
  const double* expmu_inv;
  const int* nb;          // [4][N] neighbours (up, right, down, left)
  // generic checkerboard: bonds of group g are grp_off[g] .. grp_off[g+1]-1
  int ngroups, nbonds;
  const int* grp_off;     // [ngroups + 1]
  const int* bond_sites;  // [nbonds][2] (src, trg)
  const cplx* bond_blocks;  // [4 variants][nbonds][4 blocks][2][2] row-major 2x2
  // state
  double* phi;            // (3, N, M) column-major
  cplx* mat[kNumMat];
  double* vec[kNumVec];
  cplx* ustack;           // [nelem][n*n]
  cplx* tstack;           // [nelem][n*n]
  double* dstack;         // [nelem][n]
  // time-displaced Green's functions (allocated on request): xmat holds the four
  // checkpoint chains' U and T, then Gt0[1..M], then G0t[1..M]; xvec the chains' D
  cplx* xmat;
  double* xvec;
  // factorization scratch
  double* qr_vn;          // [2n] partial column norms
  cplx* qr_cand;          // [2][W][n] candidate columns (double buffered)
  double* qr_cand_norm;   // [2][W]
  int* qr_cand_pos;       // [2][W]
  int* qr_perm;           // [n] logical -> physical column
  int* qr_lpos;           // [n] physical -> logical
  cplx* qr_tau;           // [n]
  cplx* qr_Tf;            // [16 (n + 16)] compact-WY T factors of the panels of the blocked QR
  int* inv_piv;           // [n]
  cplx* inv_bcast;        // [2][n]
  cplx* lu_small;         // [32] delta (16) and M^-1 (16) of the last single proposal
  cplx* lu_win;           // [2][4 sites][4][n] columns published between local-update rounds
  cplx* lu_diag;          // [2][N][16] diagonal site blocks published between rounds
  cplx* ld_A;             // [16][n][4] pending left factors of the delayed local updates
  cplx* ld_B;             // [16][4][n] pending right factors
  int* ld_cmd;            // [2] command and pending count of a flush round
  // local updates
  const double* uniforms; // host-drawn uniform stream (device copy)
  long long* ucursor;     // next unread uniform
  long long n_uniforms;
  double* scal;           // per-chain scalars, see ScalarId
  double* prop_log;       // [4 * log_cap] per-proposal (p_acc, u or -1, accepted, detratio imag)
  long long log_cap;
  long long* log_count;
  // group barrier
  unsigned int* bar_count;
  unsigned int* bar_gen;
  int* error;
  // per-step plan of a merged (lockstep) sweep: [steps][kPlanRoles] encoded references
  // (kind << 24 | idx, or -1), see MK_PLAN
  const int* plan;
};

// Roles of a plan row: the references one update step of a merged sweep needs per chain.
// PRE: the stack element loaded before it is rebuilt (left or right operand); ADD: source
// and destination elements of add_slice_sequence_left/right; P1/P2: the loads after it
// (P2 only at the turning points, after the element was reset to the identity); ID: the
// element reset to the identity; GW: the copy of G the drift check compares with (or none:
// no copy, no drift recorded).
enum PlanRole : int {
  PR_PRE_DU, PR_PRE_DT, PR_PRE_DD, PR_PRE_SU, PR_PRE_ST, PR_PRE_SD,
  PR_ADD_SU, PR_ADD_SD, PR_ADD_ST, PR_ADD_DU, PR_ADD_DD, PR_ADD_DT,
  PR_P1_DU, PR_P1_DT, PR_P1_DD, PR_P1_SU, PR_P1_ST, PR_P1_SD,
  PR_ID_U, PR_ID_T, PR_ID_D,
  PR_P2_DU, PR_P2_DT, PR_P2_DD, PR_P2_SU, PR_P2_ST, PR_P2_SD,
  PR_GW,
  kPlanRoles
};
__host__ __device__ inline int plan_enc(int kind, int idx) { return (kind << 24) | idx; }

// Per-chain scalar slots (ChainDesc::scal).
enum ScalarId : int {
  S_BOSON_ACTION = 0,  // p.boson_action
  S_ACC_SLICE,         // accepted proposals in the last local_updates call
  S_ACC_TOTAL,         // accepted proposals since reset
  S_PROP_TOTAL,        // proposals since reset
  S_DRIFT_LAST,        // max|G_wrapped - G_fresh| of the last recompute
  S_DRIFT_MAX,         // running max of the above
  S_LOGDET,            // log|det G| (calculate_logdet)
  S_LOGDET_A,          // log|det| scratch from inversions
  S_NONREAL,           // proposals with |Im r / Re r| > 1e-4
  S_MINRATIO,          // min |p_acc - u| over proposals that drew u
  S_DRIFT_TMP,         // accumulator of the max-abs difference kernel
  S_GB_ACTION,         // boson action before the pending global update
  S_GB_LOGDET,         // log|det G| before the pending global update (mc.s.gb_log_det)
  S_GLOBAL_ACC,        // 1 if the last global update was accepted
  S_GLOBAL_PACC,       // its acceptance probability
  S_GLOBAL_NACC,       // accepted global updates since reset
  S_GLOBAL_NPROP,      // global updates since reset
  S_LOGDET_SAVE,       // S_LOGDET kept across the unit surfaces that recompute G
  kNumScalar = 24
};

__device__ inline cplx* matptr(const ChainDesc& c, MatRef r) {
  const size_t nn = (size_t)c.n * c.n;
  if (r.kind == MK_PLAN) {
    const int e = c.plan[r.idx];
    if (e < 0) return nullptr;
    r = MatRef{e >> 24, e & 0xffffff};
  }
  if (r.kind == MK_NAMED) return c.mat[r.idx];
  if (r.kind == MK_USTACK) return c.ustack + nn * r.idx;
  if (r.kind == MK_XMAT) return c.xmat + nn * r.idx;
  return c.tstack + nn * r.idx;
}
__device__ inline double* vecptr(const ChainDesc& c, VecRef r) {
  if (r.kind == VK_PLAN) {
    const int e = c.plan[r.idx];
    if (e < 0) return nullptr;
    r = VecRef{e >> 24, e & 0xffffff};
  }
  if (r.kind == VK_NAMED) return c.vec[r.idx];
  if (r.kind == VK_XVEC) return c.xvec + (size_t)c.n * r.idx;
  return c.dstack + (size_t)c.n * r.idx;
}

__device__ inline GroupSync chain_group(const ChainDesc& c) {
  return GroupSync{c.bar_count, c.bar_gen, c.error, gridDim.x};
}

}  // namespace dqmc

// ---------------------------------------------------------------------------
// Host launch wrappers (each defined next to its kernel). All are asynchronous
// on `stream`; `chains` points to `nchains` device-resident ChainDesc entries
// and `nmax` is the largest n in the batch.
// ---------------------------------------------------------------------------
namespace dqmc {

enum BOp : int { B_LEFT = 0, B_RIGHT, B_INV_LEFT, B_INV_RIGHT, B_DAG_LEFT,
                 // checkerboard halves of effective_greens2greens!: X hop_half[B] hop_half[A] and
                 // hop_half_inv[A] hop_half_inv[B] X (no slice dependence)
                 B_EFF_RIGHT, B_EFF_INV_LEFT };

// X <- op(B_slice) X  or  X op(B_slice) with the checkerboard operators; slice is 0-based
// and the same for all chains, or taken per chain from `slices` when non-null. ngroups = 0:
// the Assaad checkerboard (plaquettes); > 0: the generic one with that many bond groups.
hipError_t launch_apply_B_chkr(hipStream_t s, const ChainDesc* chains, int nchains, int nmax, int Nmax, MatRef X,
                               BOp op, int slice, const int* slices, int ngroups = 0);
// The same with a per-chain choice between two operators of the same side (both left or
// both right): sel[k] = 0 applies op0, 1 op1, -1 leaves chain k unchanged; slices[k] per chain.
hipError_t launch_apply_B_chkr_sel(hipStream_t s, const ChainDesc* chains, int nchains, int nmax, int Nmax, MatRef X,
                                   BOp op0, BOp op1, const int* sel, const int* slices, int ngroups);
// only the eV factor: mode 0 N, 1 T, 2 C; rows = apply to rows (from the right)
hipError_t launch_apply_eV(hipStream_t s, const ChainDesc* chains, int nchains, int nmax, int Nmax, MatRef X,
                           bool rows, int mode, double power, int slice, const int* slices);

enum GemmOp : int { OP_N = 0, OP_C = 1 };
// C = diag(dl) op(A) op(B) diag(dr)  (+ beta C); dl/dr optional (kind < 0 disables)
hipError_t launch_gemm(hipStream_t s, const ChainDesc* chains, int nchains, int nmax, MatRef C,
                       MatRef A, GemmOp opA, MatRef B, GemmOp opB, VecRef dl, VecRef dr, double beta);

// elementwise helpers
hipError_t launch_copy(hipStream_t s, const ChainDesc* chains, int nchains, int nmax, MatRef dst, MatRef src);
hipError_t launch_copy_vec(hipStream_t s, const ChainDesc* chains, int nchains, int nmax, VecRef dst, VecRef src);
hipError_t launch_identity(hipStream_t s, const ChainDesc* chains, int nchains, int nmax, MatRef dst, VecRef ones);
hipError_t launch_ones_vec(hipStream_t s, const ChainDesc* chains, int nchains, int nmax, VecRef dst);
// dst = diag(d) src
hipError_t launch_scale_rows_copy(hipStream_t s, const ChainDesc* chains, int nchains, int nmax, MatRef dst,
                                  MatRef src, VecRef d);
hipError_t launch_scale_cols(hipStream_t s, const ChainDesc* chains, int nchains, int nmax, MatRef X, VecRef d);
hipError_t launch_add_diag(hipStream_t s, const ChainDesc* chains, int nchains, int nmax, MatRef X, VecRef d);
hipError_t launch_recip_vec(hipStream_t s, const ChainDesc* chains, int nchains, int nmax, VecRef dst, VecRef src);
// inv_big = 1 / max(D, 1), small = min(D, 1)
hipError_t launch_split_scales(hipStream_t s, const ChainDesc* chains, int nchains, int nmax, VecRef D,
                               VecRef inv_big, VecRef small);
hipError_t launch_adjoint(hipStream_t s, const ChainDesc* chains, int nchains, int nmax, MatRef dst, MatRef src);
// X += src diag(d)
hipError_t launch_add_scaled_cols(hipStream_t s, const ChainDesc* chains, int nchains, int nmax, MatRef X,
                                  MatRef src, VecRef d);
// X = -X
hipError_t launch_negate(hipStream_t s, const ChainDesc* chains, int nchains, int nmax, MatRef X);
// scal[S_DRIFT_LAST] = max|A - B|, scal[S_DRIFT_MAX] = max(that, old)
hipError_t launch_maxdiff(hipStream_t s, const ChainDesc* chains, int nchains, int nmax, MatRef A, MatRef B);

// A = Q diag(D) T with column-pivoted Householder QR: Q -> Qout, D -> vec, T -> Tout.
// A is consumed (overwritten by the factorization). wg = workgroups per chain.
hipError_t launch_udt(hipStream_t s, const ChainDesc* chains, int nchains, int nmax, MatRef A, VecRef D,
                      MatRef Qout, MatRef Tout, int wg);
// the two halves of launch_udt: the pivoted QR itself, then Q and T formation
hipError_t launch_qrp(hipStream_t s, const ChainDesc* chains, int nchains, int nmax, MatRef A, VecRef D, int wg);
hipError_t launch_form_T(hipStream_t s, const ChainDesc* chains, int nchains, int nmax, MatRef A, VecRef D,
                         MatRef T);
// Q of a factorisation by launch_qr_blocked (same nb), from its compact-WY panels
hipError_t launch_form_Q_blocked(hipStream_t s, const ChainDesc* chains, int nchains, int nmax, MatRef A, MatRef Q,
                                 int nb = 0);
hipError_t launch_form_TQ(hipStream_t s, const ChainDesc* chains, int nchains, int nmax, MatRef A, VecRef D,
                          MatRef Qout, MatRef Tout);

// Blocked Householder QR of A (in place, k_qrp layout; D = |Re R_kk|); prepivot = 1 orders
// the columns by decreasing initial norm first, 0 keeps them. nb: panel width, 16 (n <= 512)
// or 8; 0 picks 16 up to n = 512, else 8.
// bmap (nblocks entries): optional per-block (chain, workgroup, workgroup count) map of a 1-D
// grid, chain | w << 16 | W << 24, for workgroup counts that differ between chains
hipError_t launch_qr_blocked(hipStream_t s, const ChainDesc* chains, int nchains, int nmax, MatRef A, VecRef D,
                             int prepivot, int wg, int nb = 0, const int* bmap = nullptr, int nblocks = 0);
// B <- R^-1 Q^H B with the factors of launch_qr_blocked(prepivot = 0) in A (same nb)
hipError_t launch_qr_solve(hipStream_t s, const ChainDesc* chains, int nchains, int nmax, MatRef A, MatRef B,
                           int nb = 0);

// X <- X^-1 in place (Gauss-Jordan, partial pivoting); scal[S_LOGDET_A] = log|det X|.
hipError_t launch_inverse(hipStream_t s, const ChainDesc* chains, int nchains, int nmax, MatRef X, int wg);
// The elimination of launch_inverse alone: the inverse ends in the descriptor's M_INVT
// matrix (X holds it with permuted columns), uses inv_bcast, the barrier words and
// scal[S_LOGDET_A], and nothing else of the chain.
hipError_t launch_inverse_gather(hipStream_t s, const ChainDesc* chains, int nchains, int nmax, MatRef X, int wg);

// One local-update pass over all sites at `slice` (0-based). variant 1 updates G in global
// memory after every accepted proposal, 2 keeps each workgroup's column block of G in LDS
// (needs local_tile_fits), 3 delays the rank-4 updates and flushes them in groups (needs
// local_delayed_fits), 4 processes the sites in windows with one group barrier per window
// (needs local_window_sites(nmax, wg) >= 2: the largest window that fits in LDS; *tile
// tells whether the workgroups' own columns of G stay in LDS as well). bmap (variant 4 only):
// per-block (chain, workgroup, workgroup count) map of a 1-D grid of nblocks, from
// local_block_map; each chain then sizes its own window.
hipError_t launch_local_updates(hipStream_t s, const ChainDesc* chains, int nchains, int nmax, int slice,
                                const int* slices, double box, int wg, int log, int variant,
                                const int* bmap = nullptr, int nblocks = 0);
int local_block_map(const std::vector<int>& n, std::vector<int>& map);
bool local_tile_fits(int nmax, int wg);
bool local_delayed_fits(int nmax);
int local_window_sites(int nmax, int wg, bool* tile = nullptr);

// scal[S_LOGDET] = log|det UG| + sum log DG + log|det TG|  helper pieces
hipError_t launch_sum_log(hipStream_t s, const ChainDesc* chains, int nchains, int nmax, VecRef d, int slot,
                          int accumulate);
hipError_t launch_add_scalar(hipStream_t s, const ChainDesc* chains, int nchains, int dst, int src);
hipError_t launch_negate_scalar(hipStream_t s, const ChainDesc* chains, int nchains, int slot);

// single proposal (calc_detratio) and forced update_greens! for unit use
hipError_t launch_detratio(hipStream_t s, const ChainDesc* chains, int slice, int site, const double* new_op,
                           double* out);
hipError_t launch_force_update(hipStream_t s, const ChainDesc* chains, int nmax, int site);
// reset counters/scratch of the per-chain scalars (keeps the boson action)
hipError_t launch_reset_stats(hipStream_t s, const ChainDesc* chains, int nchains);

// boson action of the current field into scal[S_BOSON_ACTION]
hipError_t launch_boson_action(hipStream_t s, const ChainDesc* chains, int nchains, int Nmax);

hipError_t launch_boson_grad(hipStream_t s, const ChainDesc* chains, int nchains, int NMmax, double* const* grad);
hipError_t launch_daxpy(hipStream_t s, double* y, const double* x, double a, long long n);
// dS/dphi fermion part at one slice from G (slice 0-based): grad[3][N] += -Re tr[(1-G)_ii eV^-1 d eV]
// global update (global_updates.jl:1-58) of chains[0]
struct RestoreList {
  double* dst[6];
  const double* src[6];
  long long count[6];  // doubles
  int n;
};
hipError_t launch_global_shift(hipStream_t s, const ChainDesc* chain, double* phi_bak, double box);
hipError_t launch_global_decide(hipStream_t s, const ChainDesc* chain);
hipError_t launch_global_restore(hipStream_t s, const double* flag, const RestoreList& r);

hipError_t launch_grad_slice(hipStream_t s, const ChainDesc* chains, int nchains, int Nmax, int slice,
                             double* const* grad);

// Fermionic observables of one chain (fermion_measurements.jl:388-899): kind 0 pairing
// (etpc / zfpc), 1 charge density (etcdc / zfcdc), 2 current-current (zfccc). Sums over
// ntau imaginary times of Gt0 + tau sGt0 and G0t + tau sG0t (sGt0 = 0: the equal-time form on
// one G), with Gtau and G0 the equal-time Green's function, over origins and displacements;
// res[2][N] complex = the two L x L result arrays (index y + L x). tnb: [2 dirs][4 flavours]
// [N][2] hopping entries (kind 2). part: fermion_obs_blocks(ntau, N) x 2N complex scratch.
int fermion_obs_blocks(int ntau, int N);
hipError_t launch_fermion_obs(hipStream_t s, int kind, int L, int ntau, const cplx* Gtau, const cplx* G0,
                              const cplx* Gt0, size_t sGt0, const cplx* G0t, size_t sG0t, const cplx* tnb, cplx* part,
                              cplx* res);

}  // namespace dqmc
