/*
 * dqmc_hip.h -- C ABI of the MI355X DQMC sweep library (libdqmc_hip.so).
 *
 * Drop-in boundary for the DQMC sweep hot path of carstenbauer/dqmc (O(3)
 * spin-fermion model). Every entry point replaces one reference function; the
 * replaced function is cited as file:line of the reference sources. Matrices
 * cross the boundary as column-major complex128 arrays with interleaved
 * (re, im) doubles, i.e. Julia's Matrix{ComplexF64} memory layout; the field
 * phi is the reference's (3, N, M) Float64 array. All indices are 0-based.
 *
 * Conventions:
 *   - every function returns 0 on success and a negative DQMC_E* code on error;
 *     dqmc_last_error() then describes the failure (thread-local string);
 *   - host buffers are owned by the caller and only accessed during the call;
 *   - a context (one chain) or batch (many chains) must be used from one host
 *     thread at a time; different contexts are independent;
 *   - calls are asynchronous on the context's HIP stream unless they return
 *     data to the host (getters, optional out-parameters), which synchronise.
 */
#ifndef DQMC_HIP_H
#define DQMC_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

#define DQMC_OK 0
#define DQMC_E_ARG (-1)     /* invalid argument or size */
#define DQMC_E_HIP (-2)     /* HIP runtime error */
#define DQMC_E_STATE (-3)   /* call not valid in the current state */
#define DQMC_E_DEVICE (-4)  /* device-side failure (barrier timeout, uniform stream exhausted) */

typedef struct dqmc_ctx dqmc_ctx;
typedef struct dqmc_batch dqmc_batch;

/* Model parameters (reference Params, parameters.jl:4-80). The lattice is the
 * L x L periodic square lattice, N = L*L, n = 4N (flv = 4, opdim = 3). */
typedef struct {
  int L;
  int slices;       /* M = beta / delta_tau */
  int safe_mult;    /* stabilisation interval; slices % safe_mult == 0 */
  int checkerboard; /* 1: Assaad 4-site checkerboard (CBAssaad), 2: generic bond checkerboard
                       (CBGeneric, any L), 0: dense (CBFalse) */
  int edrun;        /* 1: bosonic action has only the mass term (parameters.jl:50) */
  int all_checks;   /* 1: record max|G_wrapped - G_fresh| at each recompute */
  double delta_tau, lambda, r, c, u;
} dqmc_params;

/* Host-computed hopping operators (the reference's setup code produces them:
 * init_checkerboard_matrices(_Bfield), hoppings_checkerboard.jl:65-136/199-270,
 * and init_hopping_matrix_exp(_Bfield), hoppings.jl:16-95/158-254).
 *
 * Checkerboard: chkr_hop_half[g], chkr_hop_half_inv[g], chkr_hop[g],
 * chkr_hop_inv[g] (g = A, B) are block-diagonal over the 4 flavour blocks and,
 * within a block, over disjoint 4-site plaquettes; they are passed as the list
 * of 4x4 plaquette blocks:
 *   plaq_sites[g][p][k]                      site of corner k of plaquette p of group g
 *   plaq[v][g][b][p][r][c] (complex, interleaved)  entry (r, c) of the block of
 *       variant v (0 half, 1 half_inv, 2 full, 3 full_inv), group g, flavour block b
 * expmu[4N] is the diagonal of chkr_mu = exp(+dtau mu).
 * Dense (checkerboard = 0): hop_exp / hop_exp_inv are hopping_matrix_exp(_inv)
 * as n x n complex column-major. neighbors[k][i] (k = up, right, down, left)
 * is the reference's neighbor table (lattice.jl:105-117) transposed, 0-based.
 * Generic checkerboard (checkerboard = 2; init_checkerboard_matrices(_Bfield),
 * hoppings_checkerboard_generic.jl:1-104, 199-270): l.groups as n_groups groups of disjoint
 * bonds, group g holding bonds group_offsets[g] .. group_offsets[g+1]-1 (group 0 takes the
 * full step, l.groups[1]); bond_sites[q][2] = (src, trg) of l.checkerboard[1:2, q];
 * bond_blocks[v][q][b][r][c] (complex, interleaved) = entry (r, c), r, c in {src, trg}, of the
 * bond's 2x2 block of variant v (chkr_hop_half, _half_inv, chkr_hop, _inv) in flavour block b;
 * expmu as above. The folding of chkr groups (fold_chkr_grps) is a regrouping of the same
 * product and needs no input. */
typedef struct {
  const int* plaq_sites;
  const double* plaq;
  const double* expmu;
  const double* hop_exp;
  const double* hop_exp_inv;
  const int* neighbors;
  int n_groups;
  const int* group_offsets;
  const int* bond_sites;
  const double* bond_blocks;
} dqmc_operators;

/* Counters and diagnostics of a chain. */
typedef struct {
  int current_slice;         /* 1-based like mc.s.current_slice, 0..M+1 */
  int direction;             /* +1 / -1 like mc.s.direction */
  double boson_action;       /* p.boson_action */
  double log_det;            /* mc.s.log_det (after calculate_logdet) */
  double drift_last;         /* max|G_wrapped - G_fresh| at the last recompute (stack.jl:424-429, 475-480) */
  double drift_max;          /* running maximum since the last stats reset */
  double acc_last;           /* acceptance rate of the last local_updates call */
  double accepted;           /* accepted proposals since reset */
  double proposed;           /* proposals since reset */
  double nonreal;            /* proposals with |Im r/Re r| > 1e-4 (local_updates.jl:19-20) */
  double min_ratio_gap;      /* min |p_acc - u| over proposals that drew u */
  long long uniforms_used;   /* position of the uniform stream cursor */
  int device_error;          /* nonzero after a device-side failure */
  double global_accepted;    /* accepted global updates since reset (a.acc_global) */
  double global_proposed;    /* global updates since reset (a.prop_global) */
} dqmc_stats;

const char* dqmc_last_error(void);

/* ---- single chain ------------------------------------------------------- */

/* DQMC(p) + initialize_stack (dqmc_framework.jl:105-135, stack.jl:224-242). */
int dqmc_ctx_create(const dqmc_params* p, const dqmc_operators* ops, int device, dqmc_ctx** out);
int dqmc_ctx_destroy(dqmc_ctx* ctx);
/* workgroups per chain for the cooperative kernels (QR, inverse, local updates); 0 = default */
int dqmc_ctx_set_groups(dqmc_ctx* ctx, int wg_qr, int wg_inv, int wg_local);
/* how calculate_greens (stack.jl:338-369) forms G = [1 + Ul Dl Tl (Ur Dr Tr)^H]^-1:
 * 0 (default) separates the D scales into max(D,1) and min(D,1) and inverts the
 * well-conditioned middle matrix once; 1 follows the reference verbatim (two
 * inner UDT decompositions, myrdiv and inv). Both give the same G and log|det G|
 * to rounding; the setting applies to every later recompute of the context. */
int dqmc_set_greens_route(dqmc_ctx* ctx, int route);
/* how local_updates (local_updates.jl:1-39) applies update_greens! while it sweeps the
 * sites of a slice: 1 updates G in global memory after every accepted proposal; 2 keeps
 * each workgroup's column block of G in LDS; 3 delays the rank-4 updates (up to 8 pending,
 * decisions from LDS-resident diagonal blocks) and adds them to G in one pass; 4 visits the
 * sites in windows of up to 8, keeping the window's accepted updates as factor pairs in
 * every workgroup and adding them to G once per window (one group barrier per window);
 * 0 (default) picks 4, else 2, else 3, else 1, by what fits in LDS. Same results to
 * rounding; an error when a forced variant does not fit. */
int dqmc_set_local_variant(dqmc_ctx* ctx, int variant);
/* how the stack's decompose_udt! (linalg.jl:20-39) factors inside build_stack /
 * add_slice_sequence_*: 0 (default) blocked Householder QR after pre-pivoting the
 * columns by decreasing norm (two group barriers per 16-column panel), 1 column-
 * pivoted QR with zgeqp3's pivot sequence (one barrier per column). Both are stable for
 * the stack's column-graded products and give the same G; only the stack's internal
 * U, D, T differ. Every other factorisation (dqmc_udt, the direct and time-displaced
 * Green's functions, whose chains are also row-graded) uses route 1. */
int dqmc_set_udt_route(dqmc_ctx* ctx, int route);

/* p.hsfield (parameters.jl:19): phi[a + 3*(i + N*l)] */
int dqmc_set_field(dqmc_ctx* ctx, const double* phi);
int dqmc_get_field(dqmc_ctx* ctx, double* phi);

/* The uniform stream the reference draws with rand() (local_updates.jl:9,31):
 * the device consumes it in order, 3 draws per proposal plus 1 when p_acc <= 1.
 * Uploading resets the cursor to 0. */
int dqmc_set_uniforms(dqmc_ctx* ctx, const double* u, long long count);

/* init!(mc, conf) after dqmc_set_field (dqmc_framework.jl:156-177):
 * boson action, build_stack, propagate. */
int dqmc_init(dqmc_ctx* ctx);
/* build_stack (stack.jl:251-272) */
int dqmc_build_stack(dqmc_ctx* ctx);
/* propagate (stack.jl:391-499); out-params optional (NULL), drift = -1 when no recompute */
int dqmc_propagate(dqmc_ctx* ctx, int* current_slice, int* direction, double* drift);
/* local_updates (local_updates.jl:1-39) at the current slice with box half-width `box`;
 * acc_rate optional (synchronises). */
int dqmc_local_updates(dqmc_ctx* ctx, double box, double* acc_rate);
/* `nsweeps` x M update steps (propagate + local_updates, dqmc_framework.jl:258-261 / 500-517
 * without global updates); one sweep = M steps = the reference's printed "sweep duration". */
int dqmc_sweep(dqmc_ctx* ctx, int nsweeps, double box);
int dqmc_synchronize(dqmc_ctx* ctx);

/* mc.s.greens */
int dqmc_get_greens(dqmc_ctx* ctx, double* G);
int dqmc_set_greens(dqmc_ctx* ctx, const double* G);
/* u_stack/d_stack/t_stack element k (stack.jl:87-89); any pointer may be NULL */
int dqmc_get_stack(dqmc_ctx* ctx, int k, double* U, double* D, double* T);
int dqmc_get_stats(dqmc_ctx* ctx, dqmc_stats* st);
int dqmc_reset_stats(dqmc_ctx* ctx);
/* per-proposal log (p_acc, u or -1, accepted, Im r/Re r) of the next local updates;
 * capacity in proposals, 0 disables */
int dqmc_set_proposal_log(dqmc_ctx* ctx, long long capacity);
int dqmc_get_proposal_log(dqmc_ctx* ctx, double* out, long long* count);

/* Unit surfaces (reference functions usable on host data) --------------- */
/* multiply_B_left!/right!/inv_left!/inv_right!, multiply_daggered_B_left!
 * (slice_matrices.jl:101-226 / 25-94); which = 0 left, 1 right, 2 inv_left,
 * 3 inv_right, 4 daggered_left; X is n x n, updated in place. */
int dqmc_apply_B(dqmc_ctx* ctx, int slice, int which, double* X);
/* wrap_greens (stack.jl:316-332): G <- wrapped G for (slice 1-based, direction) */
int dqmc_wrap_greens(dqmc_ctx* ctx, double* G, int slice, int direction);
/* decompose_udt (linalg.jl:10-39) of an n x n matrix */
int dqmc_udt(dqmc_ctx* ctx, const double* A, double* U, double* D, double* T);
/* No reference counterpart: the stack's default factorisation (udt route 0, blocked QR
 * after pre-pivoting by column norm) as A = U diag(D) T, for tests */
int dqmc_udt_prepivoted(dqmc_ctx* ctx, const double* A, double* U, double* D, double* T);
/* calculate_greens + calculate_logdet (stack.jl:338-385) from given UDT factors */
int dqmc_calculate_greens(dqmc_ctx* ctx, const double* Ul, const double* Dl, const double* Tl,
                          const double* Ur, const double* Dr, const double* Tr, double* G,
                          double* logdet);
/* calc_greens_and_logdet (fermion_measurements.jl:1068-1101): G(slice) from scratch */
int dqmc_greens_at(dqmc_ctx* ctx, int slice, double* G, double* logdet);
/* calc_detratio (local_updates.jl:42-59) at the current slice; any out pointer may be NULL:
 * det[2], delta[16 complex, column-major 4x4], M[16 complex] */
int dqmc_detratio(dqmc_ctx* ctx, int site, const double* new_op, double* det, double* delta, double* M);
/* update_greens! (local_updates.jl:61-95) with the delta/M of the last dqmc_detratio call */
int dqmc_update_greens(dqmc_ctx* ctx, int site);
/* mul!(C, op(A), op(B)) for n x n complex matrices (zgemm through LinearAlgebra.mul!,
 * stack.jl:291,346-366); opA/opB: 0 = none, 1 = adjoint */
int dqmc_gemm(dqmc_ctx* ctx, int opA, int opB, const double* A, const double* B, double* C);
/* inv(A) and log|det A| (Julia inv/logdet via getrf+getri, stack.jl:359, :381-383) */
int dqmc_inverse(dqmc_ctx* ctx, const double* A, double* Ainv, double* logabsdet);
/* calc_boson_action (action.jl:1-52) of the current field */
int dqmc_boson_action(dqmc_ctx* ctx, double* S);

/* No reference counterpart: dqmc_inverse for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Needs no context and allocates nothing: A, Ainv (n x n complex, column-major, interleaved)
 * and logabsdet (one double, may be NULL) are device pointers the caller owns; the
 * library's scratch (about 33 KB) is a fixed per-device array. Limit: 1 <= n <= 1024.
 * A is overwritten (it holds the inverse with permuted columns afterwards); Ainv = inv(A),
 * *logabsdet = log|det A| (NaN after a device-side failure). Asynchronous on `stream` (a
 * hipStream_t; NULL = the null stream). Calls that overlap on one device share the scratch
 * and race: order them on one stream. Same kernel and results as dqmc_inverse. */
int dqmc_inverse_noalloc(int n, double* A, double* Ainv, double* logabsdet, void* stream);

/* global_update (global_updates.jl:19-58), callable when current_slice == M and
 * direction == -1 (update, dqmc_framework.jl:507-512): shifts every component of
 * phi by -b + 2 b u with three uniforms from the stream (b = box_global), rebuilds
 * the stack, and accepts with p = exp(-dS_B) exp(log_det_old - log_det_new) (one
 * more uniform when p <= 1). On rejection field, stack, G, log_det and boson
 * action are restored. `accepted` is optional; passing it synchronises. */
int dqmc_global_update(dqmc_ctx* ctx, double box_global, int* accepted);

/* ---- measurement hand-off -------------------------------------------------- */
/* effective_greens2greens! (fermion_measurements.jl:1125-1141 checkerboard,
 * :1178-1185 dense) on an n x n host matrix, in place */
int dqmc_effective_greens2greens(dqmc_ctx* ctx, double* G);
/* allocate_tdgfs! (fermion_measurements.jl:1239-1265): device storage for Gt0/G0t
 * at every slice (2 M n^2 complex) and the four checkpoint chains (8 M/safe_mult
 * n x n complex). Needs an even M (fill_tdgf! splits at M/2 + 1). Values persist
 * between measurements, zero after allocation. */
int dqmc_tdgf_alloc(dqmc_ctx* ctx);
/* deallocate_tdgfs_stacks! (fermion_measurements.jl:1267-1284); frees Gt0/G0t too */
int dqmc_tdgf_free(dqmc_ctx* ctx);
/* measure_tdgfs! (fermion_measurements.jl:1343-1407, fill_tdgf! :1513-1541) for
 * the current field; results stay on the device */
int dqmc_measure_tdgfs(dqmc_ctx* ctx);
/* mc.s.meas.Gt0[tau], mc.s.meas.G0t[tau] (tau 1-based); either pointer may be NULL */
int dqmc_get_tdgf(dqmc_ctx* ctx, int tau, double* Gt0, double* G0t);
/* entry k (0-based) of BT0Inv (which 0), BBetaT (1), BT0 (2), BBetaTInv (3)
 * _u/_d/_t_stack (fermion_measurements.jl:1248-1259); any pointer may be NULL */
int dqmc_get_tdgf_stack(dqmc_ctx* ctx, int which, int k, double* U, double* D, double* T);
/* calc_tdgf_direct (fermion_measurements.jl:1553-1576): G(slice, 1) from scratch */
int dqmc_tdgf_direct(dqmc_ctx* ctx, int slice, double* G);

/* ---- fermionic observables ------------------------------------------------ *
 * The observables read an equal-time Green's function kept on the device (the reference's
 * etgf, measure_framework.jl:716-720) and, for the zero-frequency forms, the time-displaced
 * Green's functions of the last dqmc_measure_tdgfs. Result arrays are the reference's
 * L x L arrays [y+1, x+1] in column-major order (displacement (x, y) at index y + L x):
 * real for the pairing correlations, complex (interleaved) for the others. */
/* measure_greens (fermion_measurements.jl:225-229): calc_greens(mc, 1) from scratch and
 * effective_greens2greens!; kept as the etgf; G (host copy, n x n) optional */
int dqmc_measure_greens(dqmc_ctx* ctx, double* G);
/* the greens observable of _measure_observables! (dqmc_framework.jl:378-382): wrap_greens of
 * the chain's G at current_slice in direction +1, then effective_greens2greens!; kept as the
 * etgf (the chain's G is unchanged); G optional */
int dqmc_measure_current_greens(dqmc_ctx* ctx, double* G);
/* a host etgf (measure_framework.jl:716-717: greens loaded from disk) */
int dqmc_set_etgf(dqmc_ctx* ctx, const double* G);
/* mc.l.hopping_matrix (init_hopping_matrix, hoppings.jl:16-95 / :158-254), n x n complex:
 * the hopping amplitudes of the current operator of dqmc_measure_zfccc */
int dqmc_set_hopping_matrix(dqmc_ctx* ctx, const double* T);
/* measure_etpc! (fermion_measurements.jl:388-470) on the etgf: etpc_plus / etpc_minus */
int dqmc_measure_etpc(dqmc_ctx* ctx, double* plus, double* minus);
/* measure_etcdc! (fermion_measurements.jl:592-657) on the etgf (complex) */
int dqmc_measure_etcdc(dqmc_ctx* ctx, double* plus, double* minus);
/* measure_zfpc! (fermion_measurements.jl:500-562) on the device Gt0 */
int dqmc_measure_zfpc(dqmc_ctx* ctx, double* plus, double* minus);
/* measure_zfcdc! (fermion_measurements.jl:688-778) with greens = the etgf (complex) */
int dqmc_measure_zfcdc(dqmc_ctx* ctx, double* plus, double* minus);
/* measure_zfccc! (fermion_measurements.jl:814-899) with greens = the etgf (complex xx, yy),
 * and measure_sfdensity (:911-923) of the result (optional) */
int dqmc_measure_zfccc(dqmc_ctx* ctx, double* xx, double* yy, double* sfdensity);

/* No reference counterpart: dS/dphi for every (component, site, slice) of
 * S(phi) = S_B(phi) - log det(1 + B_M...B_1), computed from the Green's
 * functions of one stabilised pass over all slices. Also returns S (optional). */
int dqmc_grad(dqmc_ctx* ctx, double* grad, double* S);
/* No reference counterpart (its only global move is a uniform shift,
 * global_updates.jl:10-16): one Hamiltonian Monte Carlo trajectory of phi with
 * H = sum p^2/2 + S(phi). `momenta` (3, N, M) are the initial momenta (host
 * N(0,1) draws), `nsteps` leapfrog steps of size `eps` with dqmc_grad forces;
 * accepted iff dH <= 0 or u < exp(-dH) with the host uniform `u`. On rejection
 * phi is restored. The chain is left as after init! (stack rebuilt, slice M). */
int dqmc_hmc_trajectory(dqmc_ctx* ctx, const double* momenta, int nsteps, double eps, double u, int* accepted,
                        double* dH);

/* ---- batches of independent chains -------------------------------------- */
/* No reference counterpart: many chains (possibly of different L, M) advanced
 * together; chains of equal shape share every kernel launch. */
int dqmc_batch_create(const dqmc_params* p, const dqmc_operators* ops, int nchains, int device,
                      dqmc_batch** out);
int dqmc_batch_destroy(dqmc_batch* b);
int dqmc_batch_set_field(dqmc_batch* b, int chain, const double* phi);
int dqmc_batch_get_field(dqmc_batch* b, int chain, double* phi);
int dqmc_batch_set_uniforms(dqmc_batch* b, int chain, const double* u, long long count);
int dqmc_batch_init(dqmc_batch* b);
/* nsweeps sweeps of every chain; box[chain] */
int dqmc_batch_sweep(dqmc_batch* b, int nsweeps, const double* box);
int dqmc_batch_synchronize(dqmc_batch* b);
int dqmc_batch_get_greens(dqmc_batch* b, int chain, double* G);
int dqmc_batch_get_stats(dqmc_batch* b, int chain, dqmc_stats* st);
/* capture each shape group's sweep in a HIP graph and replay it (1) or launch eagerly (0).
 * No reference counterpart (the reference has no device). A sweep's kernel sequence is fixed
 * by the group's (current_slice, direction) at its start, the box and whether proposals are
 * logged, so one graph is captured per such state (two for alternating up/down sweeps) on
 * first use and replayed after; changing workgroup counts or routes drops the graphs. Same
 * results as eager launches; not used while kernel timing (dqmc_batch_profile) is on.
 * Merged sweeps (dqmc_batch_set_merged) are captured the same way, one merged sweep of every
 * chain per start state of all groups, each graph with its own plan. */
int dqmc_batch_use_graph(dqmc_batch* b, int enable);
/* No reference counterpart. 1 (default): a batch of several shape groups (any L and M, same
 * safe_mult, checkerboard kind, all_checks, box and the default calculate_greens route) is
 * swept in lockstep, every kernel launched once for all chains: all chains recompute G on the
 * same steps (M is a multiple of safe_mult), and each chain's own stack elements, slices
 * and turning points come from a per-chain plan; a chain leaves the launches when its sweeps
 * are done. Same results as sweeping the groups one by one (0) up to rounding in G (the
 * workgroup counts per launch differ). The blocked QR's and the local updates' workgroups
 * per chain follow the chain's size. */
int dqmc_batch_set_merged(dqmc_batch* b, int enable);
int dqmc_use_graph(dqmc_ctx* ctx, int enable);
/* Autotuner. No reference counterpart (the reference has one fixed CPU code path). Per
 * shape group: times the local-update kernel over (variant, workgroups per chain) and the
 * blocked QR of the stack over (workgroups per chain, panel width 16 or 8) on the group's
 * current state, `reps` timed launches each after a warm-up, and keeps the fastest. The
 * chains' G, field, scalars and uniform cursor are restored bit for bit (local updates are
 * timed on synthetic uniforms), so tuning does not change the Monte Carlo trajectory. Needs
 * an initialised batch. report (optional): 6 doubles per group -- workgroups and variant of
 * the local updates and their time in ms, workgroups and panel width of the QR and its time. */
int dqmc_batch_autotune(dqmc_batch* b, const double* box, int reps, double* report);
int dqmc_autotune(dqmc_ctx* ctx, double box, int reps, double* report);
/* dqmc_ctx_set_groups / dqmc_set_local_variant for every shape group of a batch */
int dqmc_batch_set_groups(dqmc_batch* b, int wg_qr, int wg_inv, int wg_local);
int dqmc_batch_set_local_variant(dqmc_batch* b, int variant);
/* panel width of the blocked QR (the stack's factorisations and calculate_greens): 16
 * (n <= 512) or 8; 0 restores the default, 8. No reference counterpart; results agree to
 * rounding. */
int dqmc_set_qr_panel(dqmc_ctx* ctx, int nb);
int dqmc_batch_set_qr_panel(dqmc_batch* b, int nb);

/* ---- kernel timing -------------------------------------------------------
 * No reference counterpart (the reference's TimerOutputs sections,
 * dqmc_framework.jl:36-43, time the same phases on the CPU). When a class bit
 * is set, every launch of that class is bracketed by HIP events on the chain
 * group's stream; _read synchronises, returns accumulated milliseconds and
 * launch counts per class, and clears the records. */
#define DQMC_K_APPLY_B 0  /* fused slice-matrix application (checkerboard) */
#define DQMC_K_GEMM 1     /* complex FP64 GEMM */
#define DQMC_K_QRP 2      /* column-pivoted Householder QR */
#define DQMC_K_QFORM 3    /* Q and T formation of the UDT */
#define DQMC_K_INVERSE 4  /* Gauss-Jordan inverse */
#define DQMC_K_LOCAL 5    /* local updates of one slice */
#define DQMC_K_OTHER 6    /* copies and elementwise helpers */
#define DQMC_K_NCLASS 7
int dqmc_profile(dqmc_ctx* ctx, unsigned class_mask);
int dqmc_profile_read(dqmc_ctx* ctx, double* ms, long long* count, int nclass);
int dqmc_batch_profile(dqmc_batch* b, unsigned class_mask);
int dqmc_batch_profile_read(dqmc_batch* b, double* ms, long long* count, int nclass);

#ifdef __cplusplus
}
#endif
#endif
