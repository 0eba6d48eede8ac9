/*
 * xr_hip.h -- C ABI of the MI355X (gfx950) XR hot-path library libxrhip.so.
 *
 * Plain C types only: int64_t / double / float / uint64_t and raw device
 * pointers.  Every entry point returns an int status (XR_OK on success);
 * xr_last_error() gives the message of the last failure on a context.
 *
 * Conventions shared by all entry points
 *  - A fragment has n_orbs spatial orbitals, dim = 2*n_orbs spin orbitals
 *    (alpha 0..n_orbs-1, then beta), dim <= 64.  The first n_core spatial
 *    orbitals of each spin are frozen core.
 *  - A charge block's determinants are the complete lexical enumeration of
 *    its valence occupations (general-XRCC/atomic_states/mol_fci.py:99-127);
 *    xr_configs_to_masks checks that.  State coefficients are row-major
 *    (n_states x n_configs), the layout of build_density_tensors.py:77.
 *  - Operator strings are ASCII strings over {'c','a'} (c = creation),
 *    leftmost operator first, e.g. "ccaa" = p+ q+ r s.  Supported:
 *    a c ca aa cc caa cca ccaa caaa ccca.
 *  - Dense density layout: out[((I*nJ + J)*dim^k) + idx],
 *    idx = ((p*dim + q)*dim + r)*dim + s  (general-XRCC/density_tensors.c:150).
 *  - Outputs are overwritten (the reference accumulates into zeroed storage,
 *    which gives the same result).
 *  - All pointers named d_* are device pointers on the context's device; work
 *    is queued on the context's stream and is asynchronous unless stated.
 *  - A context serialises its calls; use one context per host thread.
 */
#ifndef XR_HIP_H
#define XR_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    XR_OK = 0,
    XR_EINVAL = 1,               /* bad argument value                      */
    XR_ESHAPE = 2,               /* inconsistent sizes                      */
    XR_EINVAL_CONFIG_ORDER = 3,  /* configs not the complete lexical list   */
    XR_EPAULI = 4,               /* a config repeats a spin orbital         */
    XR_ENOMEM = 5,               /* device allocation failed                */
    XR_EHIP = 6,                 /* HIP runtime error                       */
    XR_EUNSUPPORTED = 7,         /* operator string / size not supported    */
    XR_EBUSY = 8                 /* a pinned workspace would have to grow   */
};

typedef struct xr_ctx xr_ctx;

/* Precision of the two-electron contraction GEMMs (xr_set_precision). */
enum { XR_PREC_F64 = 0, XR_PREC_SPLIT_F32 = 1, XR_PREC_SPLIT_I8 = 2 };

/* Context on `device`, queuing on `stream` (a hipStream_t; NULL = default).
 * No reference counterpart: the reference is a set of stateless C functions
 * compiled at import (general-XRCC/build_density_tensors.py:23-24). */
int xr_create(int device, void *stream, xr_ctx **out);
int xr_destroy(xr_ctx *ctx);
int xr_set_stream(xr_ctx *ctx, void *stream);
const char *xr_last_error(xr_ctx *ctx);
int xr_synchronize(xr_ctx *ctx);

/* Pins (delta = +1) or releases (delta = -1) the context's workspaces.  A
 * captured HIP graph replays the device addresses the calls had when it was
 * captured; while the pin count is positive no workspace slot is
 * reallocated, and a call that would need a larger one fails with XR_EBUSY
 * instead of freeing memory the graph still reads.  No reference
 * counterpart (the reference has no captured launches). */
int xr_pin(xr_ctx *ctx, int delta);

/* Caps the total bytes of the context's workspaces (0 = no cap): a call
 * whose scratch would exceed it fails with XR_ENOMEM instead of allocating
 * (for callers that share HBM with other work; also how the out-of-memory
 * paths are tested).  No reference counterpart. */
int xr_set_memory_limit(xr_ctx *ctx, int64_t bytes);

/* Frees the context's workspaces (they grow again on demand); XR_EBUSY while
 * pinned.  The cached excitation and side tables stay.  No reference
 * counterpart. */
int xr_trim(xr_ctx *ctx);

/* on != 0: the caller promises that every V passed to the monomer
 * two-electron adjoints on this context satisfies V[pqrs] = V[rspq] (the
 * two-body part of a Hermitian Hamiltonian), which enables their one-gather
 * path for a block with itself.  Default off.  No reference counterpart. */
int xr_assume_hermitian_v(xr_ctx *ctx, int on);

/* Kernel timing with HIP events on the context stream.  xr_profile(ctx, 1)
 * clears the records and starts bracketing every launch; xr_profile(ctx, 0)
 * stops.  xr_profile_stats sums the records of one kernel class
 * (0 FP64 MFMA tile GEMM, 1 gather, 2 expand, 3 excitation table, 4 block
 * scatter / axis packing, 5 braket assembly, 6 other (incl. split-K
 * reductions), 7 split-FP32 GEMM, 8 bandwidth-bound tall-skinny products,
 * 9 grouped FP64 tile GEMM, 10 the pair program's work-list GEMM,
 * 11 split-int8 GEMM, 12 its exponent and digit-plane passes):
 * launches, total ms, and the
 * algorithmic flops and bytes of those launches.  Synchronises the stream.
 * No reference counterpart (the reference has no timers,
 * density_tensors.c:177-178). */
int xr_profile(xr_ctx *ctx, int enable);
/* Restricts the profiler to the kernel classes whose bits are set in mask
 * (bit k = class k; default all), so that timing one class leaves the other
 * launches without the two event records per launch. */
int xr_profile_kinds(xr_ctx *ctx, uint32_t mask);
int xr_profile_stats(xr_ctx *ctx, int kind, int64_t *launches, double *total_ms, double *flops,
                     double *bytes);
/* Writes every profiled launch (kind, GEMM shape, ms, flops, bytes) as CSV. */
int xr_profile_dump(xr_ctx *ctx, const char *path);

/* Occupation lists (n_cfg x n_elec int64, ascending rows) -> 64-bit masks,
 * checking Pauli, core occupancy and that row i has lexical rank i.
 * Replaces the config marshalling of build_density_tensors.py:71-78, which
 * does no checking (density_tensors.c:49-62 silently mis-ranks).
 * Synchronous (returns the check result). */
int xr_configs_to_masks(xr_ctx *ctx, const int64_t *d_cfg, int64_t n_cfg, int n_elec,
                        int n_orbs, int n_core, uint64_t *d_masks);

/* Lexical valence rank of each mask (find_config_index,
 * general-XRCC/density_tensors.c:49-62). */
int xr_config_rank(xr_ctx *ctx, const uint64_t *d_masks, int64_t n, int n_orbs, int n_core,
                   int64_t *d_rank);

/* Excitation lists of an operator string for ket configurations [q0, q1),
 * in the reference loop order (ket config outermost, then the rightmost
 * operator's index, ..., the leftmost operator's index innermost;
 * general-XRCC/density_tensors.c:162-556).  Two passes:
 *   xr_excitation_count: d_offsets[0..q1-q0] = exclusive prefix sum of the
 *     surviving entries per ket config (d_offsets[q1-q0] = total; synchronous,
 *     *total also returned);
 *   xr_excitation_fill: entry e gets (Q, idx, P, sign) with P the bra rank.
 * n_elec_ket: electrons in the ket space. */
int xr_excitation_count(xr_ctx *ctx, const char *ops, int n_orbs, int n_core, int n_elec_ket,
                        const uint64_t *d_ket_masks, int64_t q0, int64_t q1,
                        int64_t *d_offsets, int64_t *total);
int xr_excitation_fill(xr_ctx *ctx, const char *ops, int n_orbs, int n_core, int n_elec_ket,
                       const uint64_t *d_ket_masks, int64_t q0, int64_t q1,
                       const int64_t *d_offsets, int64_t *d_q, int64_t *d_idx, int64_t *d_p,
                       int8_t *d_sign);

/* Dense transition density <I| ops |J> for all state pairs (replaces
 * a/c/ca/aa/cc/caa/cca/ccaa_tensor, general-XRCC/density_tensors.c:162-556,
 * caaa/ccca_tensor, Be-states/_attic_/deprecated/density_tensors.c:560-688,
 * including its separate bra coefficient list, and the XR order-1 strings
 * ccaaa/cccaa of the same file, :690-836).  The ket space has
 * n_elec_ket electrons and nQ = C(dim-2n_core, n_elec_ket-2n_core)
 * determinants; the bra space is implied by ops.  d_out: nI*nJ*dim^k doubles. */
int xr_density(xr_ctx *ctx, const char *ops, int n_orbs, int n_core, int n_elec_ket,
               int64_t nI, int64_t nP, const double *d_zbra,
               int64_t nJ, int64_t nQ, const double *d_zket, double *d_out);

/* Adjoint of xr_density: given Rbar = dE/d(density) in the dense layout,
 * ADDS dE/dzbra (nI x nP) into d_zbar_bra and dE/dzket (nJ x nQ) into
 * d_zbar_ket (either may be NULL).  Runs the factorised density backwards:
 * Rbar is folded onto sorted index tuples, two FP64 GEMMs give the adjoints
 * of the bra and ket gathers, and each coefficient's adjoint is collected
 * by one thread over the excitations that reach it (no atomics).  No
 * reference counterpart: it provides the dE/dz the state optimizer needs
 * (StateSpaceOptimizer/state_gradients.py:121-244). */
int xr_density_adjoint(xr_ctx *ctx, const char *ops, int n_orbs, int n_core, int n_elec_ket,
                       int64_t nI, int64_t nP, const double *d_zbra, int64_t nJ, int64_t nQ,
                       const double *d_zket, const double *d_Rbar, double *d_zbar_bra,
                       double *d_zbar_ket);

/* Monomer two-electron energy matrix E[I,J] = sum_pqrs V[p,q,r,s] <I|p+q+sr|J>
 * for one charge block, without forming the 4-index density (replaces the
 * ccaa build plus per-(I,J) monomer_2e, build_density_tensors.py:122-133 and
 * H_contractions.c:61-78).  d_V: dim^4 row-major.  d_out: nI x nJ. */
int xr_monomer_2e(xr_ctx *ctx, int n_orbs, int n_core, int n_elec,
                  int64_t nI, int64_t nP, const double *d_zbra,
                  int64_t nJ, int64_t nQ, const double *d_zket,
                  const double *d_V, double *d_out);

/* Adjoint of xr_monomer_2e for Ebar = dE/dE2 (nI x nJ): ADDS dE/dzbra,
 * dE/dzket and dE/dV (dim^4) into the given buffers (any may be NULL).
 * For a block with itself (d_zbra == d_zket) into one shared coefficient
 * adjoint it takes a one-gather path when the context was told that V is
 * Hermitian, V[pqrs] = V[rspq] (xr_assume_hermitian_v); without that
 * promise, or when it does not hold, the general two-sided path runs. */
int xr_monomer_2e_adjoint(xr_ctx *ctx, int n_orbs, int n_core, int n_elec,
                          int64_t nI, int64_t nP, const double *d_zbra,
                          int64_t nJ, int64_t nQ, const double *d_zket,
                          const double *d_V, const double *d_Ebar, double *d_zbar_bra,
                          double *d_zbar_ket, double *d_Vbar);

/* xr_monomer_2e that also leaves its intermediate W = Veff Y (Veff: V
 * antisymmetrised on sorted spin-orbital pairs, Y: the ket gathered over
 * the pairs it loses) in d_W, w_len = C(dim,2) * C(dim, n_elec-2) * nJ
 * doubles, and the adjoint that takes it instead of recomputing it (same V
 * and ket; for the gradient step that follows the energy).  The same
 * references as xr_monomer_2e (build_density_tensors.py:122-133,
 * H_contractions.c:61-78). */
int xr_monomer_2e_keep(xr_ctx *ctx, int n_orbs, int n_core, int n_elec,
                       int64_t nI, int64_t nP, const double *d_zbra,
                       int64_t nJ, int64_t nQ, const double *d_zket,
                       const double *d_V, double *d_out, double *d_W, int64_t w_len);
int xr_monomer_2e_adjoint_keep(xr_ctx *ctx, int n_orbs, int n_core, int n_elec,
                               int64_t nI, int64_t nP, const double *d_zbra,
                               int64_t nJ, int64_t nQ, const double *d_zket,
                               const double *d_V, const double *d_Ebar, double *d_zbar_bra,
                               double *d_zbar_ket, double *d_Vbar, const double *d_W, int64_t w_len);

/* Strided-batched FP64 GEMM on the matrix cores:
 *   C_b = alpha * op(A_b) * op(B_b) + beta * C_b,  b < batch.
 * transA = 0: A is M x K row-major (A[m*lda+k]); 1: A is K x M (A[k*lda+m]).
 * transB = 0: B is K x N row-major (B[k*ldb+n]); 1: B is N x K (B[n*ldb+k]).
 * C is M x N row-major.  Used by every contraction below. */
int xr_gemm_f64(xr_ctx *ctx, int transA, int transB, int64_t M, int64_t N, int64_t K,
                double alpha, const double *d_A, int64_t lda, int64_t strideA,
                const double *d_B, int64_t ldb, int64_t strideB,
                double beta, double *d_C, int64_t ldc, int64_t strideC, int64_t batch);

/* xr_gemm_f64 that allocates nothing: no reference counterpart; it is for
 * callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * caller passes A, B and C as for xr_gemm_f64; the split-K partial tiles
 * and the tall-skinny partial sums go to a static 64 MiB device array of
 * the library (one per GPU, kept until the process ends) instead of the
 * context's workspaces.  A call whose staging would exceed 64 MiB fails
 * with XR_EINVAL (the allocating entry point handles it).  Every call on a
 * GPU shares that GPU's array: calls that overlap on one GPU (other
 * streams, other contexts, other threads) race.  Results equal
 * xr_gemm_f64's (the same kernels and tile / K-split choices).  Launch
 * timing (xr_profile) is suspended for the call. */
int xr_gemm_f64_noalloc(xr_ctx *ctx, int transA, int transB, int64_t M, int64_t N, int64_t K,
                        double alpha, const double *d_A, int64_t lda, int64_t strideA,
                        const double *d_B, int64_t ldb, int64_t strideB,
                        double beta, double *d_C, int64_t ldc, int64_t strideC, int64_t batch);


/* n independent products C_q = alpha op(A_q) op(B_q) of uneven shapes
 * (host arrays of n shapes, leading dimensions and device pointers; the same
 * transposes for all), their output tiles in one grid per 24 problems of one
 * tile shape (tall-skinny products go to the single-product kernels);
 * every C_q is overwritten.  Used for the pair-term products of one XR
 * operator application (general-XRCC/hamiltonian.py:21-63, the pair terms
 * H2_mn (x) 1 of all charge groups at once). */
int xr_gemm_grouped(xr_ctx *ctx, int n, int transA, int transB, const int64_t *M, const int64_t *N,
                    const int64_t *K, double alpha, const double *const *d_A, const int64_t *lda,
                    const double *const *d_B, const int64_t *ldb, double *const *d_C,
                    const int64_t *ldc);

/* Split-FP32 GEMM: while a tile is staged, each group of 8 consecutive k
 * values of a row (column) is put on a power-of-two grid below its largest
 * value, hi = the grid multiple (11 bits), lo = fp32(x - hi); the FP32
 * matrix cores form the group's hi*hi sum exactly, which goes into an FP64
 * accumulator, and hi*lo + lo*(hi + lo) with FP32 accumulation.  Error a
 * few 2^-35 of max|a| max|b| per group.  Same arguments as xr_gemm_f64.
 * No reference counterpart. */
int xr_gemm_split_f32(xr_ctx *ctx, int transA, int transB, int64_t M, int64_t N, int64_t K,
                      double alpha, const double *d_A, int64_t lda, int64_t strideA,
                      const double *d_B, int64_t ldb, int64_t strideB,
                      double beta, double *d_C, int64_t ldc, int64_t strideC, int64_t batch);

/* Split-int8 GEMM (the Ozaki scheme on the int8 matrix cores): every row of
 * op(A) and column of op(B) is scaled by a power of two and cut into 7
 * signed 7-bit digit planes; the 28 digit-plane products of weight
 * 2^(-7 t), t <= 6, are summed exactly in int32 (K chunks of at most 65536)
 * and combined in FP64 in the epilogue.  Elementwise error at most
 * 36 K 2^-49 max|a_row| max|b_col|.  Same arguments as xr_gemm_f64.
 * No reference counterpart. */
int xr_gemm_split_i8(xr_ctx *ctx, int transA, int transB, int64_t M, int64_t N, int64_t K,
                     double alpha, const double *d_A, int64_t lda, int64_t strideA,
                     const double *d_B, int64_t ldb, int64_t strideB,
                     double beta, double *d_C, int64_t ldc, int64_t strideC, int64_t batch);

/* Vector steps of the block-Davidson solver behind xr_operator's lowest
 * roots (the matrix-free counterpart of general-XRCC/excitonic.py:56-82,
 * numpy.linalg.eigh of the assembled XR matrix), each one pass over an
 * n x b row-major block (b <= 32 for the column reductions):
 *   xr_precond_divide: Y[i,j] = X[i,j] / d, d = sign (lam[i] - theta[j]),
 *       d := +floor where |d| < floor (X == Y allowed);
 *   xr_residual: R = AX - X diag(theta) and norms[j] = ||R[:, j]|| for j < k;
 *   xr_column_norms: out[j] = sum_i A[i,j]^2 (squared != 0) or its root;
 *   xr_transpose: B[j*ldb + i] = A[i*lda + j] (rows x cols).
 * The column sums are fixed-order (bitwise reproducible).  No reference
 * counterpart (numpy does these steps inside eigh). */
int xr_precond_divide(xr_ctx *ctx, int64_t n, int64_t b, const double *d_X, int64_t ldx, double *d_Y,
                      int64_t ldy, const double *d_lam, const double *d_theta, double sign, double floor_);
int xr_residual(xr_ctx *ctx, int64_t n, int b, int k, const double *d_X, int64_t ldx, const double *d_AX,
                int64_t ldax, const double *d_theta, double *d_R, int64_t ldr, double *d_norms);
int xr_column_norms(xr_ctx *ctx, int64_t n, int b, const double *d_A, int64_t lda, double *d_out, int squared);
int xr_transpose(xr_ctx *ctx, int64_t rows, int64_t cols, const double *d_A, int64_t lda, double *d_B,
                 int64_t ldb);

/* The Davidson preconditioner's diagonal of a charge sector in the product
 * eigenbasis, and the pair terms' contribution to it:
 *   xr_sector_diag: for each of ntup tuples q (host arrays: dims[q*nf + m],
 *     off[q], device pointers e[q*nf + m] to the monomer diagonals and
 *     pd[q*np + p] to the pair diagonals, pairs (m < n) in lexicographic
 *     order, np = nf(nf-1)/2), out[off[q] + i] = sum_m e_m[j_m] +
 *     sum_(m<n) pd_mn[j_m * dims_n + j_n] for the row-major digits j of i,
 *     summed in that order; nf <= 8;
 *   xr_column_dots: out[j] = sum_i A[i,j] B[i,j] (rows x cols, fixed-order
 *     sums), the diagonal of K^T H K from K and H K.
 * No reference counterpart (the reference diagonalises densely). */
int xr_sector_diag(xr_ctx *ctx, int nf, int ntup, const int64_t *dims, const int64_t *off,
                   const double *const *d_e, const double *const *d_pd, double *d_out);
int xr_column_dots(xr_ctx *ctx, int64_t rows, int64_t cols, const double *d_A, int64_t lda, const double *d_B,
                   int64_t ldb, double *d_out);

/* Spin-orbital integral blocks cut from spatial integrals (alpha spin
 * orbitals of a fragment first), one block per list entry: kinds[i] = 1, a
 * one-body block out[(s,p),(s',q)] = d(s,s') A[o0+p, o1+q] of the n x n
 * matrix src[i] (dims[4i..4i+1] = n0, n1, offs likewise); kinds[i] = 2, a
 * two-body block out[(s0,p),(s1,q),(s2,r),(s3,s)] = 1/4 (d(s0,s2) d(s1,s3)
 * V[o0+p,o2+r,o1+q,o3+s] - d(s0,s3) d(s1,s2) V[o0+p,o3+s,o1+q,o2+r]) of the
 * n^4 chemist-notation tensor src[i].  Replaces the spin-orbital integral
 * assembly of hermitian-XRCC/get_ints.py:62-155 (orbital_solver.py:323-437)
 * for the fragment blocks the contractions read.  The adjoint adds every
 * block gradient grad[i] (null: none) into dsrc[i] in list order. */
int xr_spin_blocks(xr_ctx *ctx, int n_keys, const int *kinds, const int *dims, const int *offs,
                   const double *const *d_src, double *const *d_dst, int64_t n);
int xr_spin_blocks_adjoint(xr_ctx *ctx, int n_keys, const int *kinds, const int *dims, const int *offs,
                           const double *const *d_grad, double *const *d_dsrc, int64_t n);

/* dst = src.permute(perm) (or dst += it, accumulate != 0) for a row-major
 * 4-index src of dims[4]: dst axis a is src axis perm[a], dst row-major.
 * The permuted integral blocks of the pair adjoint and their gradients.
 * No reference counterpart (numpy transposes in the reference). */
int xr_permute4(xr_ctx *ctx, const double *d_src, const int64_t *dims, const int *perm, double *d_dst,
                int accumulate);

/* Ingredients of the certified lower bound of a charge sector's lowest XR
 * eigenvalue (pair decomposition of the operator with a Lagrange multiplier
 * on the total charge):
 *   xr_pair_dress: out[r,c] = (B[r,c] + B[c,r])/2 + wm [b_r == b_c] (hm[a_r,a_c]
 *     + hm[a_c,a_r])/2 + wn [a_r == a_c] (hn[b_r,b_c] + hn[b_c,b_r])/2, the
 *     cnt x cnt fixed-charge pair block B over pair states (pa[r], pb[r])
 *     with the shares wm, wn of the monomer matrices hm, hn;
 *   xr_lanczos: `steps` Lanczos steps (no reorthogonalisation) from a fixed
 *     pseudo-random start on each of nb <= 64 symmetric matrices (host
 *     arrays n, M of device pointers, ld): the tridiagonal entries
 *     alpha[b*steps + k], beta[b*steps + k] (device; beta = 0 after a
 *     breakdown) and anorm[b] = the largest absolute row sum of matrix b;
 *   xr_add_diagonal: M[b][i,i] += shift[b] (host array) for nb <= 64
 *     matrices.
 *   xr_cholesky_test: in-place Cholesky factorisation of nb <= 64
 *     symmetric matrices (their lower triangles; 64-column panels, the
 *     trailing updates on the FP64 matrix cores); d_info[b] = 0 when matrix b
 *     is positive definite to working precision, else the 1-based column
 *     whose pivot was not positive.  The matrices are overwritten.
 * Success on M - t I certifies t (to the factorisation's rounding
 * allowance) as a lower bound of the lowest eigenvalue.  No reference
 * counterpart (the reference diagonalises the assembled XR matrix densely,
 * general-XRCC/excitonic.py:56-82, and has no charge-sector solver). */
int xr_pair_dress(xr_ctx *ctx, int64_t cnt, const double *d_B, int64_t ldb, const int *d_pa, const int *d_pb,
                  const double *d_hm, int64_t ldm, const double *d_hn, int64_t ldn, double wm, double wn,
                  double *d_out, int64_t ldo);
int xr_lanczos(xr_ctx *ctx, int nb, const int64_t *n, const double *const *d_M, const int64_t *ld, int steps,
               double *d_alpha, double *d_beta, double *d_anorm);
int xr_add_diagonal(xr_ctx *ctx, int nb, const int64_t *n, double *const *d_M, const int64_t *ld,
                    const double *shift);
int xr_cholesky_test(xr_ctx *ctx, int nb, const int64_t *n, double *const *d_M, const int64_t *ld,
                     int *d_info);

/* Symmetric product C = alpha A A^T (trans = 0, A n x K) or alpha A^T A
 * (trans = 1, A K x n), row-major: the tiles on and above the diagonal on the
 * FP64 matrix cores, then the lower triangle mirrored (C exactly symmetric,
 * overwritten).  The pair densities of the XR gradient, sum over the other
 * fragments of v v^T (the contractions of the eigenvector with itself over
 * the other fragments' indices, StateSpaceOptimizer/state_gradients.py:193-199,
 * for every charge sum of a pair at once). */
int xr_gemm_syrk(xr_ctx *ctx, int trans, int64_t n, int64_t K, double alpha, const double *d_A, int64_t lda,
                 double *d_C, int64_t ldc);

/* The product as the two-electron contractions run it in precision `mode`
 * (XR_PREC_*): FP64; split-FP32; or in XR_PREC_SPLIT_I8 the split-int8
 * product where it pays (at least 2^33 flops, every dimension >= 128 and
 * M or N >= 1024) and FP64 elsewhere.  Same other arguments as xr_gemm_f64.
 * Used by the adjoints of those contractions.  No reference counterpart. */
int xr_gemm_mode(xr_ctx *ctx, int mode, int transA, int transB, int64_t M, int64_t N, int64_t K,
                 double alpha, const double *d_A, int64_t lda, int64_t strideA,
                 const double *d_B, int64_t ldb, int64_t strideB,
                 double beta, double *d_C, int64_t ldc, int64_t strideC, int64_t batch);

/* Selects FP64 (XR_PREC_F64, default), split-FP32 (XR_PREC_SPLIT_F32) or
 * split-int8 (XR_PREC_SPLIT_I8) GEMMs for the two-electron contractions (xr_monomer_2e, xr_dimer_2min2pls,
 * xr_dimer_1min1pls_2e, xr_dimer_ExEx, xr_trimer_*).  Densities, one-electron
 * terms and the eigenproblem stay FP64. */
int xr_set_precision(xr_ctx *ctx, int mode);

/* Batched H_contractions.c families.  Each evaluates one reference scalar
 * contraction for every pair (a, b) of state pairs at once:
 *   out[a*nB + b] = f(R1[a], R2[b], integrals)
 * with R1[a] the a-th density of fragment 1 (a = I1*nJ1 + J1, dense layout
 * above) and R2[b] the b-th density of fragment 2.  out is overwritten. */

/* sum h[p,q] R[a,p,q]   (monomer_1e / monomer_extPot, H_contractions.c:48,80);
 * out has nA entries. */
int xr_monomer_1e(xr_ctx *ctx, int64_t n, int64_t nA, const double *d_Rca, const double *d_h,
                  double *d_out);

/* sum V[p1,q1,r2,s2] Rcc1[a,p1,q1] Raa2[b,s2,r2]   (dimer_2min2pls, :95) */
int xr_dimer_2min2pls(xr_ctx *ctx, int64_t n1, int64_t n2, int64_t nA, const double *d_Rcc1,
                      int64_t nB, const double *d_Raa2, const double *d_V, double *d_out);

/* sum h[p1,q2] Rc1[a,p1] Ra2[b,q2]   (dimer_1min1pls_1e, :116) */
int xr_dimer_1min1pls_1e(xr_ctx *ctx, int64_t n1, int64_t n2, int64_t nA, const double *d_Rc1,
                         int64_t nB, const double *d_Ra2, const double *d_h, double *d_out);

/* 2 ( sum V1112[p1,q1,r1,s2] Rcca1[a,q1,p1,r1] Ra2[b,s2]
 *   + sum V1222[p1,q2,r2,s2] Rc1[a,p1] Rcaa2[b,q2,s2,r2] )   (dimer_1min1pls_2e, :129) */
int xr_dimer_1min1pls_2e(xr_ctx *ctx, int64_t n1, int64_t n2, int64_t nA, const double *d_Rc1,
                         const double *d_Rcca1, int64_t nB, const double *d_Ra2,
                         const double *d_Rcaa2, const double *d_V1112, const double *d_V1222,
                         double *d_out);

/* xr_dimer_1min1pls_1e + xr_dimer_1min1pls_2e in one pass (dimer_1min1pls_1e
 * :116 plus dimer_1min1pls_2e :129, the sum the reference forms in
 * build_H.py:56-102): out = [T | Rc1] [Ra2 | S]^T with T = Rc1 h + 2 Rcca1 W1
 * and S = 2 Rcaa2 W2^T; the output is written once. */
int xr_dimer_1min1pls(xr_ctx *ctx, int64_t n1, int64_t n2, int64_t nA, const double *d_Rc1,
                      const double *d_Rcca1, int64_t nB, const double *d_Ra2,
                      const double *d_Rcaa2, const double *d_h, const double *d_V1112,
                      const double *d_V1222, double *d_out);

/* 4 sum V[p1,q2,r1,s2] Rca1[a,p1,r1] Rca2[b,q2,s2]   (dimer_ExEx, :163) */
int xr_dimer_ExEx(xr_ctx *ctx, int64_t n1, int64_t n2, int64_t nA, const double *d_Rca1,
                  int64_t nB, const double *d_Rca2, const double *d_V, double *d_out);

/* Three-fragment families, out[(a*nB + b)*nC + c] for state pairs a, b, c of
 * fragments 1, 2, 3 (H_contractions.c:184-252):
 *   2min1pls1pls: 2 sum V[p1,q1,r2,s3] Rcc1[a,q1,p1] Ra2[b,r2] Ra3[c,s3]
 *   2pls1min1min: 2 sum V[r2,s3,p1,q1] Raa1[a,q1,p1] Rc2[b,r2] Rc3[c,s3]
 *   Ex1min1pls:   4 sum V[p1,r2,q1,s3] Rca1[a,p1,q1] Rc2[b,r2] Ra3[c,s3]   */
int xr_trimer_2min1pls1pls(xr_ctx *ctx, int64_t n1, int64_t n2, int64_t n3, int64_t nA,
                           const double *d_Rcc1, int64_t nB, const double *d_Ra2, int64_t nC,
                           const double *d_Ra3, const double *d_V, double *d_out);
int xr_trimer_2pls1min1min(xr_ctx *ctx, int64_t n1, int64_t n2, int64_t n3, int64_t nA,
                           const double *d_Raa1, int64_t nB, const double *d_Rc2, int64_t nC,
                           const double *d_Rc3, const double *d_V, double *d_out);
int xr_trimer_Ex1min1pls(xr_ctx *ctx, int64_t n1, int64_t n2, int64_t n3, int64_t nA,
                         const double *d_Rca1, int64_t nB, const double *d_Rc2, int64_t nC,
                         const double *d_Ra3, const double *d_V, double *d_out);

/* Index bookkeeping of the XR matrix (general-XRCC/test_H.py:63-108 and the
 * element layout build_H.py returns into): adds a block of 1-3 fragment
 * contractions into the tensor-product matrix,
 *   out[row * ldo + col] += scale * block[sum_k e_k * stride[k]]
 * where fragment k contributes the state pair e_k = I_k*nJ[k] + J_k,
 * row = (.. (rows_k[I_k]) * d[k+1] + ..), col likewise with cols_k[J_k].
 * diag_only[k] = 1 restricts fragment k to I_k == J_k (stride ignored), for
 * terms that carry a Kronecker delta on a spectator fragment (the extPot and
 * nuclear terms of build_H.py:85-87, 151).  d_block = NULL adds `scale`.
 * Host arrays of length nfrag; d_rows / d_cols are host arrays of device
 * pointers. */
int xr_block_scatter(xr_ctx *ctx, int nfrag, const int64_t *nI, const int64_t *nJ,
                     const int64_t *stride, const int *diag_only,
                     const int64_t *const *d_rows, const int64_t *const *d_cols,
                     const int64_t *d, const double *d_block, double scale, double *d_out,
                     int64_t ldo);

/* Tensor-product XR matrix from monomer, pair and trimer blocks for
 * n_frag <= 8 fragments with spf states each (replaces braket_loops,
 * general-XRCC/hamiltonian.py:21-83).  H1: host array of n_frag device
 * pointers (spf x spf); H2: host array of n_frag*n_frag device pointers,
 * H2[M*n_frag+N] for M < N (spf^2 x spf^2); H3: NULL or n_frag^3 pointers
 * (NULL entries skipped).  d_par: NULL for the reference assembly, or
 * n_frag*spf electron-count parities, which adds the sign of spectator
 * fragments to one-electron transfers between non-adjacent fragments
 * (the reference omits it).  d_out: spf^n_frag squared, overwritten. */
int xr_braket_assemble(xr_ctx *ctx, int n_frag, int64_t spf, const double *const *H1,
                       const double *const *H2, const double *const *H3, const int8_t *d_par,
                       double *d_out);

/* dE/d(pair block) for a dimer XR eigenvector v (d1*d2 entries):
 * G[(I1*nJ1+J1)*nB + I2*nJ2+J2] = v[r1[I1]*d2 + r2[I2]] * v[c1[J1]*d2 + c2[J2]]
 * (G[b*nA + a] when swapped), the transpose of xr_block_scatter applied to
 * v v^T without forming it. */
int xr_pair_outer(xr_ctx *ctx, const double *d_v, int64_t d2, const int64_t *d_r1,
                  const int64_t *d_c1, const int64_t *d_r2, const int64_t *d_c2, int64_t nI1,
                  int64_t nJ1, int64_t nI2, int64_t nJ2, int swapped, double *d_G);

/* GEMM autotuner (no counterpart in the reference, whose contractions are
 * fixed C loops, general-XRCC/H_contractions.c): times the four tile shapes
 * of xr_gemm_f64 with K splits 1..64 on scratch operands of the given shape
 * (reps timed launches each) against the untuned dispatch and records the
 * fastest in the context when it beats the untuned one; later xr_gemm_f64
 * calls of exactly that shape use it.  Outputs the chosen tile (0 128x128,
 * 1 128x64, 2 64x128, 3 64x64; -1 = the untuned dispatch is kept, e.g. the
 * shape takes a skinny kernel), splits and mean time.  Synchronises the
 * stream: call it outside graph capture. */
int xr_gemm_tune(xr_ctx *ctx, int transA, int transB, int64_t M, int64_t N, int64_t K,
                 int64_t batch, int reps, int *tile, int *splits, double *ms);
int xr_gemm_tune_reset(xr_ctx *ctx);

/* Matrix-free XR operator, axis packing (replaces the dense product-space
 * loops of general-XRCC/hamiltonian.py:21-63 for large fragment counts).
 * full: row-major tensor [dims[0]]...[dims[nf-1]][ncol]; packed: rows = the
 * front axes (front0 < front1, front1 = -1 for one axis) reordered through
 * d_rowmap (row r holds front index rowmap[r] = a*dims[front1] + b; NULL =
 * identity, then nrows must equal the front size), columns = the other
 * axes in order and then the column index, packed rows ld_packed apart
 * (0 = the row length).  mode 0: packed = gather(full); mode 1: full +=
 * alpha * scatter(packed).  nf <= 8. */
int xr_axes_pack(xr_ctx *ctx, int nf, const int64_t *dims, int64_t ncol, int front0, int front1,
                 const int32_t *d_rowmap, int64_t nrows, double *d_full, double *d_packed,
                 int64_t ld_packed, int mode, double alpha);

/* n_items packs (mode 0) or unpacks (mode 1) of xr_axes_pack in one launch.
 * Item q: nf[q] axes with dims[q*8 .. q*8+nf[q]-1], ncol[q], front0[q],
 * front1[q], d_rowmaps ? d_rowmaps[q] : identity, nrows[q], d_full[q],
 * d_packed[q], ld_packed[q].  The items of one call must write disjoint
 * memory. */
int xr_axes_pack_multi(xr_ctx *ctx, int n_items, const int *nf, const int64_t *dims,
                       const int64_t *ncol, const int *front0, const int *front1,
                       const int32_t *const *d_rowmaps, const int64_t *nrows, double *const *d_full,
                       double *const *d_packed, const int64_t *ld_packed, int mode, double alpha);

/* Monomer matrices along one fragment axis of many vector blocks, one launch:
 * item q is the row-major view (pre[q], n[q], post[q]) at d_src[q];
 * d_dst[q] = op(U_q) applied on the middle axis (+ beta d_dst[q]), op(U) = U
 * (trans 0) or U^T (trans 1), U_q n x n row-major, n <= 64, src != dst
 * (general-XRCC/hamiltonian.py:21-63, the monomer terms H1_m (x) 1). */
int xr_axis_apply_multi(xr_ctx *ctx, int n_items, const double *const *d_U, const int *n,
                        const int *trans, const int64_t *pre, const int64_t *post,
                        const double *const *d_src, double *const *d_dst, double beta);

/* A = A + A^T in place for a square n x n row-major matrix (the XR pair
 * matrix is symmetric for real states: the blocks of one charge combination
 * and of its transpose are built once and completed by this; no reference
 * counterpart, build_H.py:56-102 builds both). */
int xr_symmetrize(xr_ctx *ctx, double *d_A, int64_t n);

/* dst[J, I, x_k .. x_1] = src[I, J, x_1 .. x_k] (accumulate: +=) for a
 * density of k <= 5 spin-orbital indices over d <= 64 spin orbitals: the
 * density of the conjugate string between the swapped charges for real
 * states, <I|op|J> = <J|op^+|I> (as Be-states/densities.py:192-206 takes
 * the conjugates by transposition). */
int xr_conjugate(xr_ctx *ctx, const double *d_src, int64_t nI, int64_t nJ, int d, int k, double *d_dst,
                 int accumulate);

/* Pair-block adjoint input (the pair-outer layout of xr_pair_outer) read from
 * a reduced pair matrix G2 = sum over the other fragments of v v^T, whose
 * rows and columns run over the pair index a*d2 + b, reordered through
 * d_inv when given; d_G2 may hold only the diagonal block starting at row and
 * column `base` (general-XRCC/hamiltonian.py:21-63, the dE/dH2 of every
 * dimer element the loops place). */
int xr_pair_gather(xr_ctx *ctx, const double *d_G2, int64_t ldg, int64_t d2, int64_t base,
                   const int32_t *d_inv,
                   const int64_t *d_r1, const int64_t *d_c1, const int64_t *d_r2,
                   const int64_t *d_c2, int64_t nI1, int64_t nJ1, int64_t nI2, int64_t nJ2,
                   int swapped, double *d_out);

/* ---------------------------------------------------------------------------
 * Batched pair blocks (general-XRCC/build_H.py:56-102 build_matrix_elements
 * .dimer over every state pair of every charge combination of every
 * fragment pair, through the H_contractions.c:95-180 families).
 *
 * One job is one charge combination of one fragment pair.  Its family is
 * named from the charge transfer of the pair's first fragment; the family's
 * "fragment 1" is the pair's first fragment unless `transposed` (then it is
 * the second one and the block is built as [b, a]):
 *   XR_PAIR_2MIN2PLS  (two electrons move)   R1 = {cc}, R2 = {aa},
 *                     V = {V[1,1,2,2]}                     (dimer_2min2pls)
 *   XR_PAIR_1MIN1PLS  (one electron moves)   R1 = {c, cca}, R2 = {a, caa},
 *                     V = {V[1,1,1,2], V[1,2,2,2]}, h = {T + U sums [1,2]}
 *                                            (dimer_1min1pls_1e + _2e)
 *   XR_PAIR_EXEX      (no transfer)          R1 = {ca}, R2 = {ca},
 *                     V = {V[1,2,1,2]}, h = {U[1,2,2], U[2,1,1]}
 *                     (dimer_ExEx + the extPot spectator terms and `nuc`
 *                     where both fragments are diagonal; never transposed)
 * Densities are (nI nJ) x d^k row-major as xr_density writes them, over the
 * family's n1 / n2 spin orbitals.  Element (I1 J1, I2 J2) of the block (state
 * pairs of the PAIR's fragments 1 and 2), times `scale`, is written to
 *   out[(rows1[I1] * mul + rows2[I2]) * ldo + cols1[J1] * mul + cols2[J2]]
 * (index arrays on the device).  The outputs listed in the call are zeroed
 * first and symmetrised (A + A^T) last, so the jobs are the 'self' (half
 * scale) and 'source' combinations only; two jobs may not place the same
 * element.  Inputs are read-only and must stay valid while a plan lives. */
enum { XR_PAIR_2MIN2PLS = 0, XR_PAIR_1MIN1PLS = 1, XR_PAIR_EXEX = 2 };

typedef struct xr_pair_job {
    int32_t kind;
    int32_t transposed;
    int64_t n1, n2;                 /* spin orbitals of the family's fragments 1, 2 */
    int64_t nI1, nJ1, nI2, nJ2;     /* bra / ket states of the pair's fragments 1, 2 */
    const double *R1[2];
    const double *R2[2];
    const double *V[2];
    const double *h[2];
    double scale;
    double nuc;
    double *out;
    int64_t ldo;
    int64_t mul;
    const int64_t *rows1, *cols1, *rows2, *cols2;
} xr_pair_job;

typedef struct xr_pair_plan xr_pair_plan;

/* Runs the jobs once on the context's stream (intermediates in the context's
 * workspaces).  outs[i] is an out_n[i] x out_n[i] matrix with leading
 * dimension out_ld[i]; mode XR_PREC_F64, XR_PREC_SPLIT_F32 or XR_PREC_SPLIT_I8
 * (the products with two-electron integrals as split GEMMs, one launch each). */
int xr_pair_blocks(xr_ctx *ctx, const xr_pair_job *jobs, int n_jobs, double *const *outs,
                   const int64_t *out_n, const int64_t *out_ld, int n_outs, int mode);

/* The same program as a plan that owns its intermediates and device
 * descriptor tables: xr_pair_plan_run allocates nothing and launches the
 * same kernels on the same addresses every time (capturable, below). */
int xr_pair_plan_create(xr_ctx *ctx, const xr_pair_job *jobs, int n_jobs, double *const *outs,
                        const int64_t *out_n, const int64_t *out_ld, int n_outs, int mode,
                        xr_pair_plan **plan);
int xr_pair_plan_run(xr_pair_plan *plan);
/* launches, workgroups and executed FP64 flops of one run (any may be NULL) */
int xr_pair_plan_info(xr_pair_plan *plan, int64_t *launches, int64_t *workgroups, double *flops,
                      int64_t *arena_bytes);
int xr_pair_plan_destroy(xr_pair_plan *plan);

/* HIP graph capture of the context's stream without a framework: calls made
 * between xr_graph_begin and xr_graph_end are recorded (not run) and
 * xr_graph_launch replays them on the context's stream.  The context must
 * have a stream of its own (xr_create / xr_set_stream with a non-null
 * stream) and the profiler off.  The workspaces are pinned (xr_pin) from
 * begin until the graph is destroyed, so run the captured sequence once
 * before capturing it (a workspace that would have to grow fails the
 * recorded call with XR_EBUSY).  No reference counterpart: the reference
 * evaluates each element through a Python call. */
typedef struct xr_graph xr_graph;
int xr_graph_begin(xr_ctx *ctx);
int xr_graph_end(xr_ctx *ctx, xr_graph **graph);
int xr_graph_launch(xr_graph *graph);
int xr_graph_destroy(xr_graph *graph);

#ifdef __cplusplus
}
#endif
#endif
